// NCC agreement and subpixel refinement, gfx950.
//
// Arithmetic is the reference CPU backend's, operation for operation
// (include/impl/cpu/agree.hpp:27-185): sequential FP32 sums, one explicit fma
// chain per accumulator in frame order, correctly rounded division and
// square root, round-half-even for the interpolated samples. The library is
// compiled with -ffp-contract=off so no other products are fused.
// TPrec = double follows the CUDA backend's Precision::DOUBLE
// (include/impl/cuda/agree.cuh:35-65).
//
// One lane per pixel; N_EXACT > 0 fixes the stack depth at compile time,
// otherwise n <= NMAX is a runtime value and frame loops carry a uniform guard.
#pragma once

#include "device_common.hpp"

namespace bicos_hip {

__device__ __forceinline__ float psqrt(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double psqrt(double v) { return __builtin_sqrt(v); }
__device__ __forceinline__ float pfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double pfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template<int NMAX, int N_EXACT>
struct Depth {
    int rt;
    __device__ __forceinline__ int get() const { return N_EXACT > 0 ? N_EXACT : rt; }
    __device__ __forceinline__ bool has(int t) const { return N_EXACT > 0 ? t < N_EXACT : t < rt; }
};

// nxcorr (agree.hpp:27-50) on two sample vectors held as floats.
template<typename TPrec, int NMAX, int N_EXACT>
__device__ __forceinline__ TPrec nxcorr(const float (&p0)[NMAX], const float (&p1)[NMAX],
                                        Depth<NMAX, N_EXACT> n, bool has_minvar, float minvar) {
    TPrec mean0 = 0, mean1 = 0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            mean0 += (TPrec)p0[t];
            mean1 += (TPrec)p1[t];
        }
    mean0 /= (TPrec)n.get();
    mean1 /= (TPrec)n.get();

    TPrec covar = 0, var0 = 0, var1 = 0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            const TPrec d0 = (TPrec)p0[t] - mean0, d1 = (TPrec)p1[t] - mean1;
            covar = pfma(d0, d1, covar);
            var0 = pfma(d0, d0, var0);
            var1 = pfma(d1, d1, var1);
        }

    if (has_minvar && (var0 < (TPrec)minvar || var1 < (TPrec)minvar))
        return (TPrec)-1;

    return covar / psqrt(var0 * var1);
}

template<typename TIn, int NMAX, int N_EXACT>
__device__ __forceinline__ void gather(float (&p)[NMAX], const Stack& f,
                                       Depth<NMAX, N_EXACT> n, int row, int col) {
#pragma unroll
    for (int t = 0; t < NMAX; ++t)
        p[t] = n.has(t) ? (float)load_sample_cached<TIn>(f, t, row, col) : 0.0f;
}

// agree (agree.hpp:52-92): in place on the int16 disparity; every pixel of
// the corrmap is written (NaN where no correlation is evaluated).
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__device__ __forceinline__ void agree_px(int row, int col, int16_t* __restrict__ disp, int64_t dstep,
                                         const Stack& f0, const Stack& f1, int n_rt, int W,
                                         float thr, int has_minvar, float minvar,
                                         TPrec* __restrict__ corrmap, int64_t cstep) {
    const Depth<NMAX, N_EXACT> n{n_rt};

    int16_t* dp = disp + (size_t)row * dstep + col;
    const int d = *dp;
    TPrec nxc = (TPrec)NAN;
    if (d != INVALID16) {
        const int col1 = col - d;
        if (col1 < 0 || W <= col1) {
            *dp = INVALID16;
        } else {
            float p0[NMAX], p1[NMAX];
            gather<TIn>(p0, f0, n, row, col);
            gather<TIn>(p1, f1, n, row, col1);
            nxc = nxcorr<TPrec>(p0, p1, n, has_minvar != 0, minvar);
            if (nxc < (TPrec)thr)
                *dp = INVALID16;
        }
    }
    if (corrmap)
        corrmap[(size_t)row * cstep + col] = nxc;
}

template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree(int16_t* __restrict__ disp, int64_t dstep,
                                               const Stack f0, const Stack f1, int n_rt,
                                               int W, float thr, int has_minvar, float minvar,
                                               TPrec* __restrict__ corrmap, int64_t cstep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
    agree_px<TIn, TPrec, NMAX, N_EXACT>(row, col, disp, dstep, f0, f1, n_rt, W, thr, has_minvar,
                                        minvar, corrmap, cstep);
}

// k_agree over listed pixels only (dense numbering of the lists, grid-stride)
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_list(int16_t* __restrict__ disp, int64_t dstep,
                                                    const Stack f0, const Stack f1, int n_rt,
                                                    int W, float thr, int has_minvar, float minvar,
                                                    TPrec* __restrict__ corrmap, int64_t cstep,
                                                    const RowLists l) {
    const int total = l.row_start[l.H];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        int row, col;
        list_entry(l, i, row, col);
        if (disp[(size_t)row * dstep + col] == LISTED_SAME16) continue;  // keeps the windowed pass's outputs
        agree_px<TIn, TPrec, NMAX, N_EXACT>(row, col, disp, dstep, f0, f1, n_rt, W, thr, has_minvar,
                                            minvar, corrmap, cstep);
    }
}

// The grid scan of agree_subpixel (agree.hpp:160-172): NCC of the left pixel
// against the interpolated right samples at each x, first strict maximum.
// WRAP = false drops the cast of the interpolated sample (caller proves it
// is the identity).
template<typename TIn, typename TPrec, int NMAX, int N_EXACT, bool WRAP>
__device__ __forceinline__ void subpixel_scan(const float (&a)[NMAX], const float (&b)[NMAX],
                                              const float (&c)[NMAX], const TPrec (&d0)[NMAX],
                                              TPrec var0, Depth<NMAX, N_EXACT> n, float step,
                                              bool mv, float minvar, float& best_x, TPrec& best) {
    for (float x = -1.0f; x <= 1.0f; x += step) {
        float it[NMAX];
        TPrec mean1 = 0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t) {
            it[t] = 0.0f;
            if (n.has(t)) {
                const float v = a[t] * x * x + b[t] * x + c[t];
                if constexpr (WRAP)
                    it[t] = (float)(TIn)(int)__builtin_rintf(v);
                else
                    it[t] = __builtin_rintf(v);
                mean1 += (TPrec)it[t];
            }
        }
        mean1 /= (TPrec)n.get();
        TPrec covar = 0, var1 = 0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t)
            if (n.has(t)) {
                const TPrec d1 = (TPrec)it[t] - mean1;
                covar = pfma(d0[t], d1, covar);
                var1 = pfma(d1, d1, var1);
            }
        const TPrec nxc = (mv && var1 < (TPrec)minvar) ? (TPrec)-1 : covar / psqrt(var0 * var1);
        if (best < nxc) {
            best_x = x;
            best = nxc;
        }
    }
}

// agree_subpixel (agree.hpp:94-185). Writes every output pixel (NaN =
// invalid) and every corrmap pixel. The left-side mean, centred samples and
// variance do not depend on x and are computed once (same values as the
// reference's per-x recomputation).
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__device__ __forceinline__ void agree_subpixel_px(
    int row, int col, const int16_t* __restrict__ disp, int64_t dstep, const Stack& f0,
    const Stack& f1, int n_rt, int W, float thr, float step, int has_minvar, float minvar,
    float* __restrict__ out, int64_t ostep, TPrec* __restrict__ corrmap, int64_t cstep) {
    const Depth<NMAX, N_EXACT> n{n_rt};
    const bool mv = has_minvar != 0;

    const int d = disp[(size_t)row * dstep + col];
    float res = NAN;
    TPrec corr = (TPrec)NAN;
    const int col1 = col - d;

    if (d != INVALID16 && 0 <= col1 && col1 < W) {
        float p0[NMAX];
        gather<TIn>(p0, f0, n, row, col);

        if (col1 == 0 || col1 == W - 1) {
            float p1[NMAX];
            gather<TIn>(p1, f1, n, row, col1);
            corr = nxcorr<TPrec>(p0, p1, n, mv, minvar);
            if (!(corr < (TPrec)thr))
                res = (float)d;
        } else {
            float a[NMAX], b[NMAX], c[NMAX];
#pragma unroll
            for (int t = 0; t < NMAX; ++t) {
                if (n.has(t)) {
                    const float y0 = (float)load_sample_cached<TIn>(f1, t, row, col1 - 1);
                    const float y1 = (float)load_sample_cached<TIn>(f1, t, row, col1);
                    const float y2 = (float)load_sample_cached<TIn>(f1, t, row, col1 + 1);
                    a[t] = 0.5f * (y0 - 2.0f * y1 + y2);
                    b[t] = 0.5f * (-y0 + y2);
                    c[t] = y1;
                } else {
                    a[t] = b[t] = c[t] = 0.0f;
                }
            }

            TPrec mean0 = 0;
#pragma unroll
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) mean0 += (TPrec)p0[t];
            mean0 /= (TPrec)n.get();
            TPrec d0[NMAX], var0 = 0;
#pragma unroll
            for (int t = 0; t < NMAX; ++t) {
                d0[t] = 0;
                if (n.has(t)) {
                    d0[t] = (TPrec)p0[t] - mean0;
                    var0 = pfma(d0[t], d0[t], var0);
                }
            }

            // The cast (TIn)(int)rint(v) of agree.hpp:164 is the identity when
            // rint(v) lies in [0, max(TIn)]. On x in [-1, 1] the parabola
            // through (-1, y0), (0, y1), (1, y2) ranges over [min, max] of y0,
            // y2 and, when its vertex x* = -b / 2a lies inside (|b| < 2|a|),
            // the vertex value c - b^2 / 4a (magnitude below 2^15 there, so its
            // float error is < 0.01). Requiring that range to stay within
            // [1, max - 1] leaves margin for the < 0.05 float error of the
            // evaluation, so every grid sample is in range and the cast can be
            // skipped, bit-identically (and no -0.0 can arise). Waves where
            // every pixel qualifies take the cast-free scan.
            constexpr float VMAX = (float)(TIn)~(TIn)0;
            bool in_range = true;
#pragma unroll
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const float y0 = c[t] - b[t] + a[t], y2 = c[t] + b[t] + a[t];
                    float lo = fminf(y0, y2), hi = fmaxf(y0, y2);
                    if (__builtin_fabsf(b[t]) < 2.0f * __builtin_fabsf(a[t])) {
                        const float v = c[t] - b[t] * b[t] / (4.0f * a[t]);
                        lo = fminf(lo, v);
                        hi = fmaxf(hi, v);
                    }
                    in_range = in_range && lo >= 1.0f && hi <= VMAX - 1.0f;
                }

            float best_x = 0.0f;
            TPrec best = (TPrec)-1;
            if (!(mv && var0 < (TPrec)minvar)) {
                if (__all(in_range))
                    subpixel_scan<TIn, TPrec, NMAX, N_EXACT, false>(a, b, c, d0, var0, n, step, mv,
                                                                     minvar, best_x, best);
                else
                    subpixel_scan<TIn, TPrec, NMAX, N_EXACT, true>(a, b, c, d0, var0, n, step, mv,
                                                                    minvar, best_x, best);
            }
            corr = best;
            if (!(best < (TPrec)thr))
                res = (float)d - best_x;
        }
    }

    out[(size_t)row * ostep + col] = res;
    if (corrmap)
        corrmap[(size_t)row * cstep + col] = corr;
}

template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1,
    int n_rt, int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, TPrec* __restrict__ corrmap, int64_t cstep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
    agree_subpixel_px<TIn, TPrec, NMAX, N_EXACT>(row, col, disp, dstep, f0, f1, n_rt, W, thr, step,
                                                 has_minvar, minvar, out, ostep, corrmap, cstep);
}

// k_agree_subpixel over listed pixels only (as k_agree_list)
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_list(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1,
    int n_rt, int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, TPrec* __restrict__ corrmap, int64_t cstep, const RowLists l) {
    const int total = l.row_start[l.H];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        int row, col;
        list_entry(l, i, row, col);
        if (disp[(size_t)row * dstep + col] == LISTED_SAME16) continue;  // keeps the windowed pass's outputs
        agree_subpixel_px<TIn, TPrec, NMAX, N_EXACT>(row, col, disp, dstep, f0, f1, n_rt, W, thr, step,
                                                     has_minvar, minvar, out, ostep, corrmap, cstep);
    }
}

}  // namespace bicos_hip
