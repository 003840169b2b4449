// Subpixel agreement decided by the closed form: the MOMENTS_FAST precision
// mode (no reference counterpart; the stage it replaces is
// include/impl/cpu/agree.hpp:94-185).
//
// Without the rounding of the interpolated samples (agree.hpp:164), the
// correlation of the left pixel u with the parabola samples through the
// right columns y0, y1, y2 (col1 - 1, col1, col1 + 1) at grid point x is
//   cov(x) / sqrt(var0 var1(x)),
//   cov(x)  = x^2 uA + x uB + uC                  (uA, uB, uC: centred cross moments)
//   var1(x) = x^4 AA + 2 x^3 AB + x^2 (BB + 2 AC) + 2 x BC + CC   (from the 3 x 3 Gram matrix)
// One lane per pixel sums the moments exactly (FP64 sums of integer
// products), centres them exactly and scores every grid point from the
// closed form. The scan's first strict maximum, the threshold and the variance
// tests are taken on these values: there is no exact stage, so a pixel whose
// best correlation lies within the closed form's distance from the
// reference's rounded samples (about 1e-4 at 33 frames) of the threshold can
// be decided the other way than by the reference, and near-equal grid points
// can swap. The learned search window
// measures what that costs against the plain path (agreement within 0.5 px).
#pragma once

#include "device_common.hpp"

namespace bicos_hip {

// The moment sums of one pixel: FP64 sums of integer samples and products
// are exact (below 2^53), at the issue cost of FP32 (non-packed FP64 vector
// instructions issue at the FP32 rate on this part).
struct FastMoments {
    double su = 0, suu = 0, s0 = 0, s1 = 0, s2 = 0, su0 = 0, su1 = 0, su2 = 0;
    double s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0;
    __device__ __forceinline__ void add(double u, double y0, double y1, double y2) {
        su += u;
        suu = __builtin_fma(u, u, suu);
        s0 += y0; s1 += y1; s2 += y2;
        su0 = __builtin_fma(u, y0, su0); su1 = __builtin_fma(u, y1, su1);
        su2 = __builtin_fma(u, y2, su2);
        s00 = __builtin_fma(y0, y0, s00); s01 = __builtin_fma(y0, y1, s01);
        s02 = __builtin_fma(y0, y2, s02); s11 = __builtin_fma(y1, y1, s11);
        s12 = __builtin_fma(y1, y2, s12); s22 = __builtin_fma(y2, y2, s22);
    }
};

// Result of one pixel from its moments (left pixel u; right columns y0, y1,
// y2 at col1 - 1, col1, col1 + 1, clamped to the row where col1 is a border
// column): res = d - x of the best grid point, or NaN; lowvar = the left
// variance is below minvar (every match of the pixel is rejected when the
// threshold is above -1).
//
// Every centred moment is formed as an integer scaled by n (n su0 - su s0,
// ...), and the closed form's coefficients as integers scaled by 2n (cov) and
// 4n (var1): all are exact in FP64 for the supported depths and sample
// counts (below 2^53), so a right column of constant samples gives exact
// zeros. The closed form is evaluated in FP64 at each grid point (FP64 FMA
// issues at the FP32 rate here) and only the final ratio in FP32, which
// keeps the score within a few 1e-7 of the closed form's value. A grid point
// whose right variance is zero within that evaluation's rounding (below
// 2^-30 of the three columns' total variance) is not scored, as the
// reference's correlation of constant samples (0 / 0) never wins.
__device__ __forceinline__ float subpixel_fast_decide(const FastMoments& m, int n, int d, bool valid_d,
                                                      bool interior, float thr, int npts, const GridX& gx,
                                                      bool mv, float minvar, bool& lowvar) {
    const double dn = (double)n;
    const double v0 = dn * m.suu - m.su * m.su;           // n var0
    lowvar = mv && v0 < dn * (double)minvar;
    if (!valid_d) return NAN;

    const double u1 = dn * m.su1 - m.su * m.s1;           // n (centred cross moment)
    const double g11 = dn * m.s11 - m.s1 * m.s1;          // n (centred Gram entry)
    const float rv0 = __builtin_amdgcn_rsqf((float)v0);
    if (!interior) {
        // border column: the plain correlation at col1 (agree.hpp:127-141)
        const bool low1 = mv && g11 < dn * (double)minvar;
        const float corr = (lowvar || low1) ? -1.0f : (float)u1 * __builtin_amdgcn_rsqf((float)g11) * rv0;
        return !(corr < thr) ? (float)d : NAN;
    }
    if (lowvar)  // no grid point is scored: the best stays -1 at x = 0
        return !(-1.0f < thr) ? (float)d : NAN;

    const double u0 = dn * m.su0 - m.su * m.s0, u2 = dn * m.su2 - m.su * m.s2;
    const double g00 = dn * m.s00 - m.s0 * m.s0, g01 = dn * m.s01 - m.s0 * m.s1;
    const double g02 = dn * m.s02 - m.s0 * m.s2, g12 = dn * m.s12 - m.s1 * m.s2;
    const double g22 = dn * m.s22 - m.s2 * m.s2;
    // cov(x) and var1(x) of the parabola a x^2 + b x + c through the three
    // columns (a = (y0 - 2 y1 + y2) / 2, b = (y2 - y0) / 2, c = y1), as
    // 2n cov = ua x^2 + ub x + uc and 4n var1 = aa x^4 + ab2 x^3 + bbac x^2 + bc2 x + cc
    const double ua = u0 + u2 - 2.0 * u1, ub = u2 - u0, uc = 2.0 * u1;
    const double aa = g00 + g22 + 2.0 * g02 + 4.0 * g11 - 4.0 * (g01 + g12);
    const double ab2 = 2.0 * (g22 - g00) + 4.0 * (g01 - g12);
    const double bbac = g00 + g22 - 2.0 * g02 + 4.0 * (g01 + g12) - 8.0 * g11;
    const double bc2 = 4.0 * (g12 - g01);
    const double cc = 4.0 * g11;
    const double minvar4 = 4.0 * dn * (double)minvar;
    const double zero_var = 0x1p-30 * 4.0 * (g00 + g11 + g22);
    float best = -1.0f, best_x = 0.0f;
    for (int k = 0; k < npts; ++k) {
        const float xf = gx.x[k];
        const double x = (double)xf;
        const double cov = __builtin_fma(__builtin_fma(ua, x, ub), x, uc);
        const double v1 = __builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(aa, x, ab2), x, bbac), x, bc2), x, cc);
        // 2n cov / sqrt(n var0 * 4n var1) = cov / sqrt(var0 var1)
        float nxc = (float)cov * __builtin_amdgcn_rsqf((float)v1) * rv0;
        if (!(v1 > zero_var)) nxc = NAN;
        if (mv && v1 < minvar4) nxc = -1.0f;
        if (best < nxc) {  // first strict maximum (agree.hpp:168-171)
            best = nxc;
            best_x = xf;
        }
    }
    return !(best < thr) ? (float)d - best_x : NAN;
}

// The right columns the subpixel stage reads for disparity d at col: col1 =
// col - d and its neighbours, clamped to the row (the right samples of a
// pixel without a match are read at valid columns and ignored, so the
// moment loop has no branches).
struct FastCols {
    int col1, cl, cmid, cr;
    bool valid_d, interior;
    __device__ __forceinline__ FastCols(int col, int d, int W) {
        col1 = col - d;
        valid_d = d != INVALID16 && 0 <= col1 && col1 < W;
        interior = valid_d && col1 != 0 && col1 != W - 1;
        cmid = valid_d ? col1 : min(1, W - 1);
        cl = max(cmid - 1, 0);
        cr = min(cmid + 1, W - 1);
    }
};

// One pixel with its samples read from the stacks in global memory
template<typename TIn>
__device__ __forceinline__ float subpixel_fast_px(int row, int col, int d, const Stack& f0,
                                                  const Stack& f1, int n, int W, float thr, int npts,
                                                  const GridX& gx, bool mv, float minvar, bool& lowvar) {
    const FastCols c(col, d, W);
    FastMoments m;
#pragma unroll 8
    for (int t = 0; t < n; ++t)
        m.add((double)load_sample_cached<TIn>(f0, t, row, col), (double)load_sample_cached<TIn>(f1, t, row, c.cl),
              (double)load_sample_cached<TIn>(f1, t, row, c.cmid), (double)load_sample_cached<TIn>(f1, t, row, c.cr));
    return subpixel_fast_decide(m, n, d, c.valid_d, c.interior, thr, npts, gx, mv, minvar, lowvar);
}

// MOMENTS_FAST over every pixel of an int16 disparity (the plain path).
template<typename TIn>
__global__ __launch_bounds__(256) void k_subpixel_fast(const int16_t* __restrict__ disp, int64_t dstep,
                                                       const Stack f0, const Stack f1, int n_rt, int W,
                                                       float thr, int npts, const GridX gx, int has_mv,
                                                       float minvar, float* __restrict__ out,
                                                       int64_t ostep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
    bool lowvar;
    out[(size_t)row * ostep + col] = subpixel_fast_px<TIn>(
        row, col, disp[(size_t)row * dstep + col], f0, f1, n_rt, W, thr, npts, gx, has_mv != 0, minvar,
        lowvar);
}

// MOMENTS_FAST over the listed pixels of each row (entries list[row * W + i],
// i < row_count[row]; the disparity of each is read from disp at its
// column). One lane per pixel; block (row, b) takes entries b * 256 + j *
// 256 * gridDim.y, j = 0, 1, ..., so that a row listing many pixels spreads
// over several blocks instead of looping in one.
template<typename TIn>
__global__ __launch_bounds__(256) void k_subpixel_fast_list(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt, int W,
    float thr, int npts, const GridX gx, int has_mv, float minvar, float* __restrict__ out,
    int64_t ostep, const int16_t* __restrict__ list, const int* __restrict__ row_count) {
    const int row = blockIdx.x, count = row_count[row];
    for (int i = blockIdx.y * 256 + threadIdx.x; i < count; i += gridDim.y * 256) {
        const int col = list[(size_t)row * W + i];
        const int d = disp[(size_t)row * dstep + col];
        if (d == INVALID16 || d == LISTED_SAME16) continue;  // output NaN already (the pixel is listed)
        bool lowvar;
        out[(size_t)row * ostep + col] = subpixel_fast_px<TIn>(
            row, col, d, f0, f1, n_rt, W, thr, npts, gx, has_mv != 0, minvar, lowvar);
    }
}

}  // namespace bicos_hip
