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
__device__ __forceinline__ void gather(float (&p)[NMAX], const FrameTable& f,
                                       Depth<NMAX, N_EXACT> n, int row, int col) {
#pragma unroll
    for (int t = 0; t < NMAX; ++t)
        p[t] = n.has(t) ? (float)load_sample_cached<TIn>(f, t, row, col) : 0.0f;
}

// agree (agree.hpp:52-92): in place on the int16 disparity; every pixel of
// the corrmap is written (NaN where no correlation is evaluated).
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree(int16_t* __restrict__ disp, int64_t dstep,
                                               const FrameTable f0, const FrameTable f1, int n_rt,
                                               int W, float thr, int has_minvar, float minvar,
                                               TPrec* __restrict__ corrmap, int64_t cstep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
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

// agree_subpixel (agree.hpp:94-185). Writes every output pixel (NaN =
// invalid) and every corrmap pixel. The left-side mean, centred samples and
// variance do not depend on x and are computed once (same values as the
// reference's per-x recomputation).
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel(
    const int16_t* __restrict__ disp, int64_t dstep, const FrameTable f0, const FrameTable f1,
    int n_rt, int W, float thr, float step, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, TPrec* __restrict__ corrmap, int64_t cstep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
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

            float best_x = 0.0f;
            TPrec best = (TPrec)-1;
            if (!(mv && var0 < (TPrec)minvar)) {
                for (float x = -1.0f; x <= 1.0f; x += step) {
                    float it[NMAX];
                    TPrec mean1 = 0;
#pragma unroll
                    for (int t = 0; t < NMAX; ++t) {
                        it[t] = 0.0f;
                        if (n.has(t)) {
                            const float v = a[t] * x * x + b[t] * x + c[t];
                            it[t] = (float)(TIn)(int)__builtin_rintf(v);
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
                    const TPrec nxc = (mv && var1 < (TPrec)minvar) ? (TPrec)-1
                                                                   : covar / psqrt(var0 * var1);
                    if (best < nxc) {
                        best_x = x;
                        best = nxc;
                    }
                }
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

}  // namespace bicos_hip
