// Subpixel agreement in the HALF_MFMA precision mode (no reference
// counterpart; the plain mode is k_agree_subpixel).
//
// Without the rounding of the interpolated samples (agree.hpp:164), the
// subpixel correlation at grid point x is a closed-form function of a few
// dot products. With u = left samples minus their mean and y'_j = right
// column j minus its mean (j = col1-1, col1, col1+1):
//   a' = (y'_0 - 2 y'_1 + y'_2) / 2,  b' = (y'_2 - y'_0) / 2,  c' = y'_1
//   covar(x) = x^2 u.a' + x u.b' + u.c'
//   var1(x)  = x^4 a'.a' + 2 x^3 a'.b' + x^2 (b'.b' + 2 a'.c') + 2 x b'.c' + c'.c'
//   nxc(x)   = covar(x) / sqrt(var0 var1(x))
// so each grid point costs O(1) instead of O(n). The three cross products
// u.y'_j come from the matrix cores: a wave takes 2 tiles of 32 consecutive
// left pixels; for each tile the centred left samples (32 x n) and the
// centred samples of a 64-column right window (n x 64), both in FP16 scaled
// by 2^-8, are contracted with v_mfma_f32_32x32x16_f16 (FP32 accumulation),
// and each pixel picks its three entries. Pixels whose three columns fall
// outside the window take the same products on the vector ALU. Left means
// and variances and right means and variances are summed in FP64. The
// winner x is then rescored exactly (the plain path's FP32 arithmetic with
// rounded samples), which gives the correlation that is written to the
// corrmap and tested against the threshold.
#pragma once

#include "agree.hpp"
#include "device_common.hpp"

namespace bicos_hip {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr int HALF_WIN = 64;   // right window columns per 32-pixel tile

template<int KP>
struct HalfLds {
    _Float16 u[64][KP];          // centred left samples of the wave's 64 pixels, x 2^-8
    _Float16 y[HALF_WIN][KP];    // centred right window samples, x 2^-8
    float d[32][HALF_WIN];       // u . y' for the current tile
    float r0[HALF_WIN], r1[HALF_WIN], r2[HALF_WIN];
};

// exact FP32 rescoring of one grid point (agree.hpp:162-166 + nxcorr)
template<typename TIn, int NMAX, int N_EXACT>
__device__ __forceinline__ float rescore(const float (&p0)[NMAX], const Stack& f1, int row,
                                         int col1, float x, Depth<NMAX, N_EXACT> n, bool mv,
                                         float minvar) {
    float it[NMAX];
#pragma unroll
    for (int t = 0; t < NMAX; ++t) {
        it[t] = 0.0f;
        if (n.has(t)) {
            const float y0 = (float)load_sample_cached<TIn>(f1, t, row, col1 - 1);
            const float y1 = (float)load_sample_cached<TIn>(f1, t, row, col1);
            const float y2 = (float)load_sample_cached<TIn>(f1, t, row, col1 + 1);
            const float a = 0.5f * (y0 - 2.0f * y1 + y2), b = 0.5f * (-y0 + y2), c = y1;
            it[t] = (float)(TIn)(int)__builtin_rintf(a * x * x + b * x + c);
        }
    }
    return nxcorr<float>(p0, it, n, mv, minvar);
}

template<typename TIn, int NMAX, int N_EXACT, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_agree_subpixel_half(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1,
    int n_rt, int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, float* __restrict__ corrmap, int64_t cstep) {
    constexpr int KP = (NMAX + 15) / 16 * 16;   // frames padded to the MFMA K step
    constexpr int KS = KP / 16;
    extern __shared__ uint4 lds_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    HalfLds<KP>& L = ((HalfLds<KP>*)lds_raw)[wave];
    const Depth<NMAX, N_EXACT> n{n_rt};
    const bool mv = has_minvar != 0;
    const int row = blockIdx.y;
    const int col = (blockIdx.x * WAVES + wave) * 64 + lane;

    // ---- phase 0: the pixel of this lane
    int d = INVALID16;
    if (col < W) d = disp[(size_t)row * dstep + col];
    const int col1 = col - d;
    const bool valid = col < W && d != INVALID16 && 0 <= col1 && col1 < W;
    const bool border = valid && (col1 == 0 || col1 == W - 1);
    const bool interior = valid && !border;

    float p0[NMAX];
    gather<TIn>(p0, f0, n, row, valid ? col : 0);
    double sum0 = 0.0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t) sum0 += (double)p0[t];
    const double mean0 = sum0 / (double)n.get();
    double var0d = 0.0;
#pragma unroll
    for (int t = 0; t < KP; ++t) {
        const double c = (t < NMAX && n.has(t)) ? (double)p0[t] - mean0 : 0.0;
        var0d += c * c;
        L.u[lane][t] = (_Float16)((float)c * (1.0f / 256.0f));
    }
    const float var0 = (float)var0d;
    const float mean0f = (float)mean0;

    float G[3] = {0.0f, 0.0f, 0.0f}, M00 = 0, M01 = 0, M02 = 0, M11 = 0, M12 = 0, M22 = 0;
    bool have = false;

    for (int T = 0; T < 2; ++T) {
        // window of this tile: [lowest col1 - 1, +64) over its interior pixels
        const bool mine = interior && (lane >> 5) == T;
        int lo = mine ? col1 - 1 : 0x7FFFFFFF;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) lo = min(lo, __shfl_xor(lo, o));
        lo = __builtin_amdgcn_readlane(lo, 32 * T);
        if (lo == 0x7FFFFFFF)
            continue;
        const int ws = lo;

        // stage window column ws + lane
        const int cw = ws + lane;
        {
            float yc[NMAX];
            gather<TIn>(yc, f1, n, row, cw < W ? cw : W - 1);
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < NMAX; ++t) s += (double)yc[t];
            const double m = s / (double)n.get();
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < NMAX; ++t) {
                const double c = n.has(t) ? (double)yc[t] - m : 0.0;
                v += c * c;
                yc[t] = (float)c;
            }
            float r1 = 0.0f, r2 = 0.0f;
#pragma unroll
            for (int t = 0; t < NMAX; ++t) {
                r1 = __builtin_fmaf(yc[t], __shfl_down(yc[t], 1), r1);
                r2 = __builtin_fmaf(yc[t], __shfl_down(yc[t], 2), r2);
            }
            L.r0[lane] = (float)v;
            L.r1[lane] = r1;
            L.r2[lane] = r2;
#pragma unroll
            for (int t = 0; t < KP; ++t)
                L.y[lane][t] = (_Float16)(t < NMAX ? yc[t] * (1.0f / 256.0f) : 0.0f);
        }
        __syncthreads();

        // d[m][c] = u_(32T + m) . y'_(ws + c), two 32 x 32 output tiles
        {
            const int nl = lane & 31, g = lane >> 5;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float16v acc = {0};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const half8 a = *(const half8*)&L.u[32 * T + nl][16 * s + 8 * g];
                    const half8 b = *(const half8*)&L.y[32 * j + nl][16 * s + 8 * g];
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    L.d[8 * (i / 4) + 4 * g + (i % 4)][32 * j + nl] = acc[i];
            }
        }
        __syncthreads();

        if (mine && col1 + 1 - ws < HALF_WIN) {
            const int c = col1 - 1 - ws, m = lane - 32 * T;
            G[0] = L.d[m][c] * 65536.0f;
            G[1] = L.d[m][c + 1] * 65536.0f;
            G[2] = L.d[m][c + 2] * 65536.0f;
            M00 = L.r0[c]; M01 = L.r1[c]; M02 = L.r2[c];
            M11 = L.r0[c + 1]; M12 = L.r1[c + 1]; M22 = L.r0[c + 2];
            have = true;
        }
        __syncthreads();
    }

    if (interior && !have) {
        // right columns outside the window: the same products on the vector
        // ALU, streaming the three columns twice (means, then products)
        double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t)
            if (n.has(t)) {
                s0 += (double)load_sample_cached<TIn>(f1, t, row, col1 - 1);
                s1 += (double)load_sample_cached<TIn>(f1, t, row, col1);
                s2 += (double)load_sample_cached<TIn>(f1, t, row, col1 + 1);
            }
        const float m0 = (float)(s0 / n.get()), m1 = (float)(s1 / n.get()), m2 = (float)(s2 / n.get());
        double m00 = 0, m11 = 0, m22 = 0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t)
            if (n.has(t)) {
                const float y0 = (float)load_sample_cached<TIn>(f1, t, row, col1 - 1) - m0;
                const float y1 = (float)load_sample_cached<TIn>(f1, t, row, col1) - m1;
                const float y2 = (float)load_sample_cached<TIn>(f1, t, row, col1 + 1) - m2;
                const float uu = p0[t] - mean0f;
                G[0] = __builtin_fmaf(uu, y0, G[0]);
                G[1] = __builtin_fmaf(uu, y1, G[1]);
                G[2] = __builtin_fmaf(uu, y2, G[2]);
                m00 += (double)y0 * y0;
                m11 += (double)y1 * y1;
                m22 += (double)y2 * y2;
                M01 = __builtin_fmaf(y0, y1, M01);
                M02 = __builtin_fmaf(y0, y2, M02);
                M12 = __builtin_fmaf(y1, y2, M12);
            }
        M00 = (float)m00; M11 = (float)m11; M22 = (float)m22;
    }

    float res = NAN, corr = NAN;
    if (border) {
        float p1[NMAX];
        gather<TIn>(p1, f1, n, row, col1);
        corr = nxcorr<float>(p0, p1, n, mv, minvar);
        if (!(corr < thr)) res = (float)d;
    } else if (interior) {
        const float ua = 0.5f * (G[0] - 2.0f * G[1] + G[2]), ub = 0.5f * (G[2] - G[0]), uc = G[1];
        const float aa = 0.25f * (M00 + 4.0f * M11 + M22 - 4.0f * M01 + 2.0f * M02 - 4.0f * M12);
        const float ab = 0.25f * (M22 - M00 - 2.0f * M12 + 2.0f * M01);
        const float bb = 0.25f * (M00 - 2.0f * M02 + M22);
        const float ac = 0.5f * (M01 - 2.0f * M11 + M12);
        const float bc = 0.5f * (M12 - M01);
        const float cc = M11;
        // approximate scan: best and runner-up grid points
        float best_x = 0.0f, best = -1.0f, second_x = 0.0f, second = -1.0f;
        if (!(mv && var0 < minvar)) {
            for (float x = -1.0f; x <= 1.0f; x += step) {
                const float cov = (ua * x + ub) * x + uc;
                const float v1 = (((aa * x + 2.0f * ab) * x + (bb + 2.0f * ac)) * x + 2.0f * bc) * x + cc;
                const float nxc = (mv && v1 < minvar) ? -1.0f : cov * __builtin_amdgcn_rsqf(var0 * v1);
                if (best < nxc) {
                    second = best;
                    second_x = best_x;
                    best = nxc;
                    best_x = x;
                } else if (second < nxc) {
                    second = nxc;
                    second_x = x;
                }
            }
        }
        // winners rescored exactly; the runner-up too when the approximate
        // scores are too close to rank (first strict maximum in grid order)
        corr = best > -1.0f ? rescore<TIn>(p0, f1, row, col1, best_x, n, mv, minvar) : -1.0f;
        if (second > -1.0f && best - second < 2e-3f) {
            const float c2 = rescore<TIn>(p0, f1, row, col1, second_x, n, mv, minvar);
            if (c2 > corr || (c2 == corr && second_x < best_x)) {
                corr = c2;
                best_x = second_x;
            }
        }
        if (!(corr < thr)) res = (float)d - best_x;
    }
    if (col < W) {
        out[(size_t)row * ostep + col] = res;
        if (corrmap) corrmap[(size_t)row * cstep + col] = corr;
    }
}

}  // namespace bicos_hip
