// Subpixel agreement in the MOMENTS precision mode (no reference
// counterpart; the plain mode is k_agree_subpixel, agree.hpp).
//
// Without the rounding of the interpolated samples (agree.hpp:164) the
// subpixel correlation is a closed-form function of the moments of the left
// pixel u and the three right columns y_0, y_1, y_2 (col1 - 1 .. col1 + 1):
// with centred cross moments uY_j and Gram matrix G_jk,
//   covar(x) = x^2 uA + x uB + uC,   uA = (uY0 - 2 uY1 + uY2) / 2, uB = (uY2 - uY0) / 2, uC = uY1
//   var1(x)  = x^4 AA + 2 x^3 AB + x^2 (BB + 2 AC) + 2 x BC + CC   (from G)
// One lane per pixel accumulates the integer-valued moments in FP64 (exact:
// sums of products of 16-bit samples stay below 2^53), evaluates the closed
// form in FP32 at every grid point and bounds what the left-out rounding can
// change there. Where the bound settles accept/reject, the pixel is done:
// the reference's decision, with the closed-form winner as the grid point
// (which can differ from the exact scan's where two grid points within half
// a pixel of each other lie within the bound; a farther rival sends the
// pixel to the exact stage).
// Every other pixel is marked pending and gets the exact FP32 subpixel
// stage (k_agree_subpixel_list) afterwards, so accept/reject is the
// reference's for every pixel. Samples are streamed from memory (twice
// where the reference's two-pass sums need them), not held in registers.
#pragma once

#include "agree.hpp"
#include "device_common.hpp"

namespace bicos_hip {

// The reference's nxcorr (agree.hpp:27-50) on the left pixel (col) and the
// right samples it(t): means, then fma chains in frame order, streaming the
// samples twice instead of holding them. Identical arithmetic to nxcorr().
template<typename TIn, int NMAX, int N_EXACT, typename Right>
__device__ __forceinline__ float nxcorr_stream(const Stack& f0, int row, int col, Right&& it,
                                               Depth<NMAX, N_EXACT> n, bool mv, float minvar) {
    float mean0 = 0.0f, mean1 = 0.0f;
#pragma unroll 3
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            mean0 += (float)load_sample_cached<TIn>(f0, t, row, col);
            mean1 += it(t);
        }
    mean0 /= (float)n.get();
    mean1 /= (float)n.get();
    float covar = 0.0f, var0 = 0.0f, var1 = 0.0f;
#pragma unroll 3
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            const float d0 = (float)load_sample_cached<TIn>(f0, t, row, col) - mean0, d1 = it(t) - mean1;
            covar = __builtin_fmaf(d0, d1, covar);
            var0 = __builtin_fmaf(d0, d0, var0);
            var1 = __builtin_fmaf(d1, d1, var1);
        }
    if (mv && (var0 < minvar || var1 < minvar))
        return -1.0f;
    return covar / psqrt(var0 * var1);
}

// Status bits of the MOMENTS stage, per pixel
constexpr uint8_t ST_PENDING = 1;   // the exact stage decides this pixel
constexpr uint8_t ST_HOPELESS = 2;  // left variance rejects every match

template<typename TIn, int NMAX, int N_EXACT>
__device__ __forceinline__ void agree_subpixel_moments_px(
    int row, int col, const int16_t* __restrict__ disp, int64_t dstep, const Stack& f0,
    const Stack& f1, int n_rt, int W, float thr, float step, int has_minvar, float minvar,
    float* __restrict__ out, int64_t ostep, uint8_t* __restrict__ status) {
    const Depth<NMAX, N_EXACT> n{n_rt};
    const bool mv = has_minvar != 0;
    const int d = disp[(size_t)row * dstep + col];
    const int col1 = col - d;
    const bool valid_d = d != INVALID16 && 0 <= col1 && col1 < W;
    const bool interior = valid_d && col1 != 0 && col1 != W - 1;
    const int c1 = interior ? col1 : 1;  // in range for the loads of other lanes
    float res = NAN;
    uint8_t st = 0;

    // One pass: moments of the left samples (every pixel: its variance
    // decides the fallback; FP64, exact for integer samples) and, for
    // interior matches, of the three right columns, with the reach of the
    // parabolas. The cross moments are FP32 sums of samples shifted by the
    // first frame's (products below 2^32; relative error < 1e-5 of the
    // centred moments, covered by the bound's slack below).
    double su = 0, suu = 0;
    float s0 = 0, s1 = 0, s2 = 0, su0 = 0, su1 = 0, su2 = 0;
    float s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0, su_f = 0;
    float ymin = 3.0e38f, ymax = -3.0e38f;
    const float u_ref = (float)load_sample_cached<TIn>(f0, 0, row, col);
    float y_ref0 = 0.0f, y_ref1 = 0.0f, y_ref2 = 0.0f;
    if (interior) {
        y_ref0 = (float)load_sample_cached<TIn>(f1, 0, row, c1 - 1);
        y_ref1 = (float)load_sample_cached<TIn>(f1, 0, row, c1);
        y_ref2 = (float)load_sample_cached<TIn>(f1, 0, row, c1 + 1);
    }
    // 11 frames per round: 44 loads in flight per lane
#pragma unroll 11
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            const float uf = (float)load_sample_cached<TIn>(f0, t, row, col);
            const double u = uf;
            su += u;
            suu = __builtin_fma(u, u, suu);
            if (interior) {
                const float f0y = (float)load_sample_cached<TIn>(f1, t, row, c1 - 1);
                const float f1y = (float)load_sample_cached<TIn>(f1, t, row, c1);
                const float f2y = (float)load_sample_cached<TIn>(f1, t, row, c1 + 1);
                // reach of the frame's parabola on [-1, 1]: the end values and,
                // where it is convex (concave) with the vertex inside, the
                // vertex, which lies above (below) y1 -/+ |y2 - y0| / 4
                const float r4 = 0.25f * fabsf(f2y - f0y);
                ymin = fminf(ymin, fminf(fminf(f0y, f2y), f1y - r4));
                ymax = fmaxf(ymax, fmaxf(fmaxf(f0y, f2y), f1y + r4));
                const float us = uf - u_ref, y0 = f0y - y_ref0, y1 = f1y - y_ref1, y2 = f2y - y_ref2;
                su_f += us;
                s0 += y0; s1 += y1; s2 += y2;
                su0 = __builtin_fmaf(us, y0, su0); su1 = __builtin_fmaf(us, y1, su1);
                su2 = __builtin_fmaf(us, y2, su2);
                s00 = __builtin_fmaf(y0, y0, s00); s01 = __builtin_fmaf(y0, y1, s01);
                s02 = __builtin_fmaf(y0, y2, s02); s11 = __builtin_fmaf(y1, y1, s11);
                s12 = __builtin_fmaf(y1, y2, s12); s22 = __builtin_fmaf(y2, y2, s22);
            }
        }
    const double inv = 1.0 / (double)n.get();
    const double var0d = suu - su * su * inv;
    // The reference tests its FP32 two-pass variance against min_variance
    // (agree.hpp:158); the exact value decides unless it is within the FP32
    // error of the threshold, where the two-pass value is formed.
    bool lowvar = false;
    if (mv) {
        const double mvd = (double)minvar, slack = 1e-4 * var0d + 1.0;
        if (var0d < mvd - slack) {
            lowvar = true;
        } else if (var0d <= mvd + slack) {
            float mean0 = 0.0f;
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) mean0 += (float)load_sample_cached<TIn>(f0, t, row, col);
            mean0 /= (float)n.get();
            float var0 = 0.0f;
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const float c = (float)load_sample_cached<TIn>(f0, t, row, col) - mean0;
                    var0 = __builtin_fmaf(c, c, var0);
                }
            lowvar = var0 < minvar;
        }
    }
    if (lowvar && thr > -1.0f) st |= ST_HOPELESS;
    if (valid_d) {
        if (!interior) {
            const float corr = nxcorr_stream<TIn>(f0, row, col, [&](int t) {
                return (float)load_sample_cached<TIn>(f1, t, row, col1);
            }, n, mv, minvar);
            if (!(corr < thr)) res = (float)d;
        } else if (lowvar) {
            if (!(-1.0f < thr)) res = (float)d;  // no grid point is scored (agree.hpp:158)
        } else {
            // centred moments (FP64 from the FP32 sums)
            const double su_s = su_f, d0 = s0, d1 = s1, d2 = s2;
            const double u0 = su0 - su_s * d0 * inv, u1 = su1 - su_s * d1 * inv, u2 = su2 - su_s * d2 * inv;
            const double g00 = s00 - d0 * d0 * inv, g01 = s01 - d0 * d1 * inv, g02 = s02 - d0 * d2 * inv;
            const double g11 = s11 - d1 * d1 * inv, g12 = s12 - d1 * d2 * inv, g22 = s22 - d2 * d2 * inv;
            const float ua = (float)(0.5 * (u0 - 2.0 * u1 + u2)), ub = (float)(0.5 * (u2 - u0)), uc = (float)u1;
            const float aa = (float)(0.25 * (g00 + 4.0 * g11 + g22 - 4.0 * g01 + 2.0 * g02 - 4.0 * g12));
            const float ab = (float)(0.25 * (g22 - g00 - 2.0 * g12 + 2.0 * g01));
            const float bb = (float)(0.25 * (g00 - 2.0 * g02 + g22));
            const float ac = (float)(0.5 * (g01 - 2.0 * g11 + g12));
            const float bc = (float)(0.5 * (g12 - g01));
            const float cc = (float)g11;
            // Bound on what the left-out rounding can change. If every
            // frame's parabola stays inside [1, max(TIn) - 1] (with 0.1 for
            // the FP32 evaluation and the approximate vertex), the cast of
            // agree.hpp:164 is the identity and every interpolated sample
            // moves by at most 0.51 (0.5 plus the FP32 evaluation error):
            // |e| <= delta = 0.51 sqrt(n). Then with s = sqrt(var1),
            // covariance and norm move by at most sqrt(var0) delta and delta,
            // the correlation by at most 2 delta / (s - delta) (+1e-4 for the
            // FP32 moments and arithmetic of both forms and the exact versus
            // two-pass left variance), and var1 itself by at most
            // 2 s delta + delta^2 (+1e-4 var1).
            constexpr float VMAX = (float)(TIn)~(TIn)0;
            const float delta = 0.51f * __builtin_sqrtf((float)n.get());
            float best_x = 0.0f, best = -1.0f, eps_best = 2.0f;
            float upper = -1.0f;  // max over x of closed form + bound
            bool settled = ymin >= 1.1f && ymax <= VMAX - 1.1f;  // (NaN-safe: NaN fails)
            const float rv0 = __builtin_amdgcn_rsqf((float)var0d);
            const auto bound = [&](float sv) {
                return 2.000001f * delta * __builtin_amdgcn_rcpf(sv - delta) + 1e-4f;
            };
            for (float x = -1.0f; x <= 1.0f; x += step) {
                const float cov = (ua * x + ub) * x + uc;
                const float v1 = (((aa * x + 2.0f * ab) * x + (bb + 2.0f * ac)) * x + 2.0f * bc) * x + cc;
                const float r1 = __builtin_amdgcn_rsqf(v1);
                const float sv = v1 * r1;  // sqrt(var1)
                const float margin = 2.0f * sv * delta + delta * delta + 1e-4f * v1 + 1.0f;
                float nxc = -1.0f, eps = 0.0f;
                if (mv && v1 < minvar) {
                    // below the variance test: also exactly when below by the margin
                    if (v1 >= minvar - margin) settled = false;
                } else {
                    if (mv && v1 < minvar + margin) settled = false;
                    if (!(sv > 2.0f * delta)) settled = false;  // no useful bound here
                    nxc = cov * r1 * rv0;
                    eps = bound(sv);
                }
                upper = fmaxf(upper, nxc + eps);
                if (best < nxc) {
                    best = nxc;
                    best_x = x;
                    eps_best = eps;
                }
            }
            // Accepted with the closed-form grid point when the bound settles
            // it and no grid point half a pixel or more away could beat the
            // winner; else the exact stage.
            if (settled && best - eps_best >= thr) {
                float far = -1.0f;
                for (float x = -1.0f; x <= 1.0f; x += step) {
                    if (fabsf(x - best_x) < 0.5f) continue;
                    const float cov = (ua * x + ub) * x + uc;
                    const float v1 = (((aa * x + 2.0f * ab) * x + (bb + 2.0f * ac)) * x + 2.0f * bc) * x + cc;
                    const float r1 = __builtin_amdgcn_rsqf(v1);
                    if (mv && v1 < minvar) continue;  // settled: exactly -1 there
                    far = fmaxf(far, cov * r1 * rv0 + bound(v1 * r1));
                }
                if (far < best - eps_best)
                    res = (float)d - best_x;
                else
                    st |= ST_PENDING;
            } else if (!(settled && upper < thr)) {
                st |= ST_PENDING;
            }
        }
    }
    out[(size_t)row * ostep + col] = res;
    status[(size_t)row * W + col] = st;
}

template<typename TIn, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_moments(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, uint8_t* __restrict__ status) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
    agree_subpixel_moments_px<TIn, NMAX, N_EXACT>(row, col, disp, dstep, f0, f1, n_rt, W, thr, step,
                                                  has_minvar, minvar, out, ostep, status);
}

// over listed pixels only (row r = blockIdx.x: row_count[r] columns at list + r * W)
template<typename TIn, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_moments_list(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, uint8_t* __restrict__ status, const int16_t* __restrict__ list,
    const int* __restrict__ row_count) {
    const int row = blockIdx.x, i = blockIdx.y * 256 + threadIdx.x;
    if (i >= row_count[row])
        return;
    agree_subpixel_moments_px<TIn, NMAX, N_EXACT>(row, list[(size_t)row * W + i], disp, dstep, f0, f1,
                                                  n_rt, W, thr, step, has_minvar, minvar, out, ostep,
                                                  status);
}

// The exact subpixel stage (agree.hpp:94-185 arithmetic, bit-identical to
// k_agree_subpixel) for listed pixels, one pixel per 32 lanes: lane j scores
// grid points j, j + 32, ... (x_k formed by the reference's running float
// sum) and the lanes reduce to the reference's first strict maximum. Clears
// ST_PENDING of the pixels it decides (status nullable).
template<typename TIn, int NMAX, int N_EXACT>
__device__ __forceinline__ void agree_subpixel_wide_px(
    int row, int col, int j, const int16_t* __restrict__ disp, int64_t dstep, const Stack& f0,
    const Stack& f1, int n_rt, int W, float thr, float step, int npts, int has_minvar,
    float minvar, float* __restrict__ out, int64_t ostep, float* __restrict__ corrmap,
    int64_t cstep, uint8_t* __restrict__ status) {
    const Depth<NMAX, N_EXACT> n{n_rt};
    const bool mv = has_minvar != 0;
    const int d = disp[(size_t)row * dstep + col];
    const int col1 = col - d;
    float res = NAN, corr = NAN;
    if (d != INVALID16 && 0 <= col1 && col1 < W) {
        if (col1 == 0 || col1 == W - 1) {
            corr = nxcorr_stream<TIn>(f0, row, col, [&](int t) {
                return (float)load_sample_cached<TIn>(f1, t, row, col1);
            }, n, mv, minvar);
            if (!(corr < thr)) res = (float)d;
        } else {
            // samples held in registers (one load round), the parabola
            // coefficients, left mean, centred samples and variance formed
            // as k_agree_subpixel forms them
            float p0[NMAX], a[NMAX], b[NMAX], c[NMAX];
#pragma unroll
            for (int t = 0; t < NMAX; ++t) {
                p0[t] = a[t] = b[t] = c[t] = 0.0f;
                if (n.has(t)) {
                    p0[t] = (float)load_sample_cached<TIn>(f0, t, row, col);
                    const float y0 = (float)load_sample_cached<TIn>(f1, t, row, col1 - 1);
                    const float y1 = (float)load_sample_cached<TIn>(f1, t, row, col1);
                    const float y2 = (float)load_sample_cached<TIn>(f1, t, row, col1 + 1);
                    a[t] = 0.5f * (y0 - 2.0f * y1 + y2);
                    b[t] = 0.5f * (-y0 + y2);
                    c[t] = y1;
                }
            }
            float mean0 = 0.0f;
#pragma unroll
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) mean0 += p0[t];
            mean0 /= (float)n.get();
            float var0 = 0.0f;
#pragma unroll
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    p0[t] -= mean0;  // centred from here on
                    var0 = __builtin_fmaf(p0[t], p0[t], var0);
                }
            float best = -1.0f;
            int best_k = 0x7FFFFFFF;  // none: x stays 0 (agree.hpp:159)
            if (!(mv && var0 < minvar)) {
                // x_j by the reference's running sum, then 32 steps per round
                float x = -1.0f;
                for (int m = 0; m < j; ++m) x += step;
                for (int k = j; k < npts; k += 32) {
                    float it[NMAX], mean1 = 0.0f;
#pragma unroll
                    for (int t = 0; t < NMAX; ++t) {
                        it[t] = 0.0f;
                        if (n.has(t)) {
                            it[t] = (float)(TIn)(int)__builtin_rintf(a[t] * x * x + b[t] * x + c[t]);
                            mean1 += it[t];
                        }
                    }
                    mean1 /= (float)n.get();
                    float covar = 0.0f, var1 = 0.0f;
#pragma unroll
                    for (int t = 0; t < NMAX; ++t)
                        if (n.has(t)) {
                            const float d1 = it[t] - mean1;
                            covar = __builtin_fmaf(p0[t], d1, covar);
                            var1 = __builtin_fmaf(d1, d1, var1);
                        }
                    const float nxc = (mv && var1 < minvar) ? -1.0f : covar / psqrt(var0 * var1);
                    if (best < nxc) {  // lane-local first strict maximum
                        best = nxc;
                        best_k = k;
                    }
                    for (int m = 0; m < 32; ++m) x += step;
                }
            }
            // reduce over the pixel's 32 lanes: larger value, then smaller k
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) {
                const float ob = __shfl_xor(best, o, 32);
                const int ok = __shfl_xor(best_k, o, 32);
                if (ob > best || (ob == best && ok < best_k)) {
                    best = ob;
                    best_k = ok;
                }
            }
            float bx = 0.0f;
            if (best_k != 0x7FFFFFFF) {
                bx = -1.0f;
                for (int k = 0; k < best_k; ++k) bx += step;
            }
            corr = best;
            if (!(best < thr)) res = (float)d - bx;
        }
    }
    if (j == 0) {
        out[(size_t)row * ostep + col] = res;
        if (corrmap) corrmap[(size_t)row * cstep + col] = corr;
        if (status) status[(size_t)row * W + col] &= (uint8_t)~ST_PENDING;
    }
}


template<typename TIn, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_wide(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int npts, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, float* __restrict__ corrmap, int64_t cstep, const int16_t* __restrict__ list,
    const int* __restrict__ row_count, uint8_t* __restrict__ status) {
    const int row = blockIdx.x, j = threadIdx.x & 31;
    const int count = row_count[row];
    for (int i = blockIdx.y * 8 + (threadIdx.x >> 5); i < count; i += gridDim.y * 8)  // uniform per pixel
        agree_subpixel_wide_px<TIn, NMAX, N_EXACT>(row, list[(size_t)row * W + i], j, disp, dstep, f0, f1,
                                                   n_rt, W, thr, step, npts, has_minvar, minvar, out,
                                                   ostep, corrmap, cstep, status);
}

// Per-row lists (row_count[r] columns at list + r * W, any order; row_count
// zeroed before) of the pixels with a status bit in `pending_mask`, or
// (pending_mask == 0) of the fallback set: invalid result and not hopeless.
// The row is blockIdx.x so that rows spread over the XCDs.
__global__ __launch_bounds__(256) void k_status_list(const uint8_t* __restrict__ status,
                                                     const float* __restrict__ res, int64_t rstep,
                                                     int pending_mask, int W,
                                                     int16_t* __restrict__ list,
                                                     int* __restrict__ row_count) {
    __shared__ int wsum[4], base;
    const int row = blockIdx.x, col = blockIdx.y * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool take = false;
    if (col < W) {
        const uint8_t st = status[(size_t)row * W + col];
        take = pending_mask ? (st & pending_mask) != 0
                            : (__builtin_isnan(res[(size_t)row * rstep + col]) && !(st & ST_HOPELESS));
    }
    const uint64_t m = __ballot(take);
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        base = total ? atomicAdd(&row_count[row], total) : 0;
    }
    __syncthreads();
    if (take) {
        int off = __popcll(m & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) off += wsum[w];
        list[(size_t)row * W + base + off] = (int16_t)col;
    }
}

__global__ void k_zero_i32(int* p, int count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0;
}

}  // namespace bicos_hip
