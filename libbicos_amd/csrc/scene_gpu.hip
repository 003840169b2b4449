// Device form of the synthetic scene generator (csrc/scene.cpp): the same
// construction and the same hashed random numbers, evaluated on the GPU so
// that training the learned window on thousands of full-size scenes does not
// wait on the host. The double-precision arithmetic follows the host code
// operation by operation (no contraction); the two differ only where the
// device and host libm round sin/cos/log differently in the last place.
#include <hip/hip_runtime.h>

#include <limits>

#include "host_common.hpp"
#include "scene_field.hpp"

using namespace bicos_hip;
using namespace bicos_scene;

namespace bicos_hip {
namespace {

struct Taps {
    double g[2 * BLUR_R + 1];
};

constexpr int MM_BLOCKS = 256;  // partial min/max blocks per segment

__global__ __launch_bounds__(256) void k_scene_field(const Field f, int W, int H, double* __restrict__ raw) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int y = (int)(i / W), x = (int)(i % W);
    raw[i] = field_value(f, (double)x, (double)y);
}

// per segment (blockIdx.y) of `count` values: MM_BLOCKS partial (min, max)
template<typename T>
__global__ __launch_bounds__(256) void k_minmax_partial(const T* __restrict__ x, size_t count,
                                                        T* __restrict__ part) {
    __shared__ T smin[256], smax[256];
    const T* seg = x + (size_t)blockIdx.y * count;
    T lo = std::numeric_limits<T>::max(), hi = std::numeric_limits<T>::lowest();
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)MM_BLOCKS * 256) {
        const T v = seg[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smin[threadIdx.x] = smin[threadIdx.x + o] < smin[threadIdx.x] ? smin[threadIdx.x + o] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + o] > smax[threadIdx.x] ? smax[threadIdx.x + o] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((size_t)blockIdx.y * MM_BLOCKS + blockIdx.x) * 2] = smin[0];
        part[((size_t)blockIdx.y * MM_BLOCKS + blockIdx.x) * 2 + 1] = smax[0];
    }
}

template<typename T>
__global__ __launch_bounds__(256) void k_minmax_final(const T* __restrict__ part, T* __restrict__ mm) {
    __shared__ T smin[MM_BLOCKS], smax[MM_BLOCKS];
    const T* p = part + (size_t)blockIdx.x * MM_BLOCKS * 2;
    smin[threadIdx.x] = p[threadIdx.x * 2];
    smax[threadIdx.x] = p[threadIdx.x * 2 + 1];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smin[threadIdx.x] = smin[threadIdx.x + o] < smin[threadIdx.x] ? smin[threadIdx.x + o] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + o] > smax[threadIdx.x] ? smax[threadIdx.x + o] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mm[blockIdx.x * 2] = smin[0];
        mm[blockIdx.x * 2 + 1] = smax[0];
    }
}

// disparity field affinely normalised to [dmin, dmax]
__global__ __launch_bounds__(256) void k_scene_disp(const double* __restrict__ raw, const double* __restrict__ mm,
                                                    size_t count, double dmin, double dmax,
                                                    float* __restrict__ D) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const double fmin = mm[0], fmax = mm[1];
    const double scale = fmax > fmin ? (dmax - dmin) / (fmax - fmin) : 0.0;
    D[i] = (float)(dmin + (raw[i] - fmin) * scale);
}

// dark left pixels and their right correspondences (darkR cleared before)
__global__ __launch_bounds__(256) void k_scene_dark(uint64_t seed, const float* __restrict__ D, int W, int H,
                                                    double p_dark, uint8_t* __restrict__ darkL,
                                                    uint8_t* __restrict__ darkR) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int y = (int)(i / W), x = (int)(i % W);
    const bool dark = u01d(hash5(seed, S_DARK, 0, (uint64_t)y, (uint64_t)x)) < p_dark;
    darkL[i] = dark ? 1 : 0;
    if (dark) {
        const long xr = (long)rint((double)x - (double)D[i]);  // lrint, half to even
        if (xr >= 0 && xr < W) darkR[(size_t)y * W + xr] = 1;
    }
}

// left coordinate seen by each right pixel: one lane per row (warp_row)
__global__ __launch_bounds__(64) void k_scene_src(const float* __restrict__ D, int W, int H,
                                                  double right_only, float* __restrict__ src,
                                                  double* __restrict__ zbuf) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    if (y >= H) return;
    warp_row(D + (size_t)y * W, W, right_only, src + (size_t)y * W, zbuf + (size_t)y * W);
}

__global__ __launch_bounds__(256) void k_scene_noise(uint64_t seed, int H, int CW, float* __restrict__ noise) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= (size_t)H * CW) return;
    const int y = (int)(i / CW), x = (int)(i % CW);
    noise[(size_t)t * H * CW + i] = u01(hash5(seed, S_SPECKLE, (uint64_t)t, (uint64_t)y, (uint64_t)x));
}

__global__ __launch_bounds__(256) void k_scene_hblur(const float* __restrict__ in, const Taps g, int H, int CW,
                                                     float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)H * CW) return;
    const size_t base = (size_t)blockIdx.y * H * CW;
    const int y = (int)(i / CW), x = (int)(i % CW);
    double s = 0.0;
    for (int k = -BLUR_R; k <= BLUR_R; ++k) {
        const int xx = min(max(x + k, 0), CW - 1);
        s += g.g[k + BLUR_R] * in[base + (size_t)y * CW + xx];
    }
    out[base + i] = (float)s;
}

__global__ __launch_bounds__(256) void k_scene_vblur(const float* __restrict__ in, const Taps g, int H, int CW,
                                                     float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)H * CW) return;
    const size_t base = (size_t)blockIdx.y * H * CW;
    const int y = (int)(i / CW), x = (int)(i % CW);
    double s = 0.0;
    for (int k = -BLUR_R; k <= BLUR_R; ++k) {
        const int yy = min(max(y + k, 0), H - 1);
        s += g.g[k + BLUR_R] * in[base + (size_t)yy * CW + x];
    }
    out[base + i] = (float)s;
}

template<typename T>
__global__ __launch_bounds__(256) void k_scene_views(const float* __restrict__ P, const float* __restrict__ pmm,
                                                     const float* __restrict__ src,
                                                     const uint8_t* __restrict__ darkL,
                                                     const uint8_t* __restrict__ darkR, uint64_t seed,
                                                     int W, int H, int CW, double lo, double hi,
                                                     double sigma, double dark_value,
                                                     T* __restrict__ left, T* __restrict__ right) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= (size_t)W * H) return;
    const int y = (int)(i / W), x = (int)(i % W);
    const float pmin = pmm[2 * t], pmax = pmm[2 * t + 1];
    const double k = pmax > pmin ? (hi - lo) / (double)(pmax - pmin) : 0.0;
    const float* Pt = P + (size_t)t * H * CW + (size_t)y * CW;
    const double vmax = (double)std::numeric_limits<T>::max();
    auto scaled = [&](double xs) {
        xs = fmin(fmax(xs, 0.0), (double)(CW - 1));
        const int x0 = (int)floor(xs);
        const int x1 = min(x0 + 1, CW - 1);
        const double w = xs - x0;
        const double v = (1.0 - w) * Pt[x0] + w * Pt[x1];
        return lo + (v - pmin) * k;
    };
    auto quantise = [&](double v) {
        v = rint(v);
        return (T)fmin(fmax(v, 0.0), vmax);
    };
    T* Lt = left + (size_t)t * W * H;
    T* Rt = right + (size_t)t * W * H;
    if (darkL[i])
        Lt[i] = (T)dark_value;
    else
        Lt[i] = quantise(scaled(x) + sigma * gauss_noise(hash5(seed, S_NOISE_L, (uint64_t)t, (uint64_t)y, (uint64_t)x)));
    if (darkR[i])
        Rt[i] = (T)dark_value;
    else
        Rt[i] = quantise(scaled((double)src[i]) + sigma * gauss_noise(hash5(seed, S_NOISE_R, (uint64_t)t, (uint64_t)y, (uint64_t)x)));
}

}  // namespace
}  // namespace bicos_hip

extern "C" int bicos_hip_scene_generate(int32_t W, int32_t H, int32_t n, int32_t depth, double dmin,
                                        double dmax, uint64_t seed, double lo, double hi,
                                        double sigma, double p_dark, double dark_value,
                                        void* left, void* right, float* disp, void* stream) {
    if (W <= 0 || H <= 0 || n <= 0 || !(dmax >= dmin) || !left || !right || (depth != 1 && depth != 2))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "bad scene parameters");
    hipStream_t s = (hipStream_t)stream;
    const Field f = make_field(seed, W, H);
    const int CW0 = W + (int)std::ceil(dmax) + 8;  // canvas columns left pixels can see
    const int CW = CW0 + W;                          // + a region only disoccluded right pixels see
    Taps taps;
    blur_taps(taps.g);

    const size_t npx = (size_t)W * H, ncv = (size_t)H * CW;
    // raw: the field values, later the warp's depth buffer
    const size_t b_raw = align_up(npx * 8, 256), b_D = align_up(npx * 4, 256),
                 b_dark = align_up(npx, 256), b_src = align_up(npx * 4, 256),
                 b_cv = align_up((size_t)n * ncv * 4, 256),
                 b_part = align_up((size_t)std::max(n, 1) * MM_BLOCKS * 2 * 8, 256),
                 b_mm = align_up((size_t)std::max(n, 1) * 2 * 8, 256);
    int dev, rc;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws;
    if ((rc = ws_acquire(dev, b_raw + b_D + 2 * b_dark + b_src + 2 * b_cv + b_part + b_mm, s, &ws, WS_SCENE)))
        return rc;
    uint8_t* p = (uint8_t*)ws;
    double* raw = (double*)p; p += b_raw;
    float* D = (float*)p; p += b_D;
    uint8_t* darkL = p; p += b_dark;
    uint8_t* darkR = p; p += b_dark;
    float* src = (float*)p; p += b_src;
    float* cv0 = (float*)p; p += b_cv;
    float* cv1 = (float*)p; p += b_cv;
    void* part = p; p += b_part;
    void* mm = p;

    auto body = [&]() -> int {
        const dim3 gpx((unsigned)((npx + 255) / 256)), gcv((unsigned)((ncv + 255) / 256), n), blk(256);
        LAUNCH("scene", k_scene_field, gpx, blk, 0, s, f, W, H, raw);
        LAUNCH("scene", k_minmax_partial<double>, dim3(MM_BLOCKS, 1), blk, 0, s, (const double*)raw, npx, (double*)part);
        LAUNCH("scene", k_minmax_final<double>, dim3(1), blk, 0, s, (const double*)part, (double*)mm);
        LAUNCH("scene", k_scene_disp, gpx, blk, 0, s, (const double*)raw, (const double*)mm, npx, dmin, dmax, D);
        HIP_TRY(hipMemsetAsync(darkR, 0, npx, s));
        LAUNCH("scene", k_scene_dark, gpx, blk, 0, s, seed, (const float*)D, W, H, p_dark, darkL, darkR);
        LAUNCH("scene", k_scene_src, dim3((H + 63) / 64), dim3(64), 0, s, (const float*)D, W, H, (double)CW0, src, raw);
        LAUNCH("scene", k_scene_noise, gcv, blk, 0, s, seed, H, CW, cv0);
        LAUNCH("scene", k_scene_hblur, gcv, blk, 0, s, (const float*)cv0, taps, H, CW, cv1);
        LAUNCH("scene", k_scene_vblur, gcv, blk, 0, s, (const float*)cv1, taps, H, CW, cv0);
        LAUNCH("scene", k_minmax_partial<float>, dim3(MM_BLOCKS, n), blk, 0, s, (const float*)cv0, ncv, (float*)part);
        LAUNCH("scene", k_minmax_final<float>, dim3(n), blk, 0, s, (const float*)part, (float*)mm);
        if (depth == 1)
            LAUNCH("scene", k_scene_views<uint8_t>, dim3((unsigned)((npx + 255) / 256), n), blk, 0, s,
                   (const float*)cv0, (const float*)mm, (const float*)src, (const uint8_t*)darkL,
                   (const uint8_t*)darkR, seed, W, H, CW, lo, hi, sigma, dark_value, (uint8_t*)left, (uint8_t*)right);
        else
            LAUNCH("scene", k_scene_views<uint16_t>, dim3((unsigned)((npx + 255) / 256), n), blk, 0, s,
                   (const float*)cv0, (const float*)mm, (const float*)src, (const uint8_t*)darkL,
                   (const uint8_t*)darkR, seed, W, H, CW, lo, hi, sigma, dark_value, (uint16_t*)left, (uint16_t*)right);
        if (disp) HIP_TRY(hipMemcpyAsync(disp, D, npx * 4, hipMemcpyDeviceToDevice, s));
        return BICOS_OK;
    };
    rc = body();
    const int rc2 = ws_release(dev, s, WS_SCENE);
    return rc ? rc : rc2;
}
