// Subpixel agreement from closed-form moments: the MOMENTS precision mode
// (no reference counterpart) and the bit-exact pruned form of the plain
// SINGLE stage (k_agree_subpixel_pruned; the full scan is k_agree_subpixel,
// agree.hpp).
//
// Without the rounding of the interpolated samples (agree.hpp:164) the
// subpixel correlation is a closed-form function of the moments of the left
// pixel u and the three right columns y_0, y_1, y_2 (col1 - 1 .. col1 + 1):
// with centred cross moments uY_j and Gram matrix G_jk,
//   covar(x) = x^2 uA + x uB + uC,   uA = (uY0 - 2 uY1 + uY2) / 2, uB = (uY2 - uY0) / 2, uC = uY1
//   var1(x)  = x^4 AA + 2 x^3 AB + x^2 (BB + 2 AC) + 2 x BC + CC   (from G)
// One lane per pixel accumulates the moments (left moments in FP64; cross
// moments as FP32 sums of samples shifted by the first frame's in MOMENTS,
// exact FP64 sums in the pruned form), evaluates the closed form in FP32 at
// every grid point and bounds what the left-out rounding and the FP32
// arithmetic can change there.
//
// MOMENTS: where the bound settles accept/reject, the pixel is done: the
// reference's decision, with the closed-form winner as the grid point (which
// can differ from the exact scan's where two grid points within half a pixel
// of each other lie within the bound; a farther rival sends the pixel to the
// exact stage). Every other pixel is pending: before the kernel writes its
// results, the wave runs the exact FP32 subpixel stage on its pending pixels
// (exact_pending_wave) while their samples are still in the L2 cache; so
// accept/reject is the reference's for every pixel.
//
// Pruned (bit-exact): the bound picks the grid points that can hold the
// scan's maximum; a single one that clears the threshold is the result, else
// exact_candidates_wave scores the candidates with the reference's FP32
// arithmetic.
#pragma once

#include <type_traits>

#include "agree.hpp"
#include "device_common.hpp"

namespace bicos_hip {

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The reference's nxcorr (agree.hpp:27-50; nxcorrd, agree.cuh:35-65, with
// TPrec = double) on the left pixel (col) and the right samples it(t):
// means, then fma chains in frame order, streaming the samples twice instead
// of holding them. Identical arithmetic to nxcorr().
template<typename TIn, typename TPrec, int NMAX, int N_EXACT, typename Right>
__device__ __forceinline__ TPrec nxcorr_stream(const Stack& f0, int row, int col, Right&& it,
                                               Depth<NMAX, N_EXACT> n, bool mv, float minvar) {
    TPrec mean0 = 0, mean1 = 0;
#pragma unroll 3
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            mean0 += (TPrec)load_sample_cached<TIn>(f0, t, row, col);
            mean1 += (TPrec)it(t);
        }
    mean0 /= (TPrec)n.get();
    mean1 /= (TPrec)n.get();
    TPrec covar = 0, var0 = 0, var1 = 0;
#pragma unroll 3
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            const TPrec d0 = (TPrec)load_sample_cached<TIn>(f0, t, row, col) - mean0,
                        d1 = (TPrec)it(t) - mean1;
            covar = pfma(d0, d1, covar);
            var0 = pfma(d0, d0, var0);
            var1 = pfma(d1, d1, var1);
        }
    if (mv && (var0 < (TPrec)minvar || var1 < (TPrec)minvar))
        return (TPrec)-1;
    return covar / psqrt(var0 * var1);
}

// Status bits of the MOMENTS stage, per pixel
constexpr uint8_t ST_PENDING = 1;   // the exact stage decides this pixel (cleared by it)
constexpr uint8_t ST_HOPELESS = 2;  // left variance rejects every match

// Pending pixels the exact stage takes per round of one wave, and the LDS
// it stages their samples in: per wave EXACT_GROUPS * NMAX entries of
// (left sample, parabola coefficients a, b, c) of frame t.
constexpr int EXACT_GROUPS = 4;

// The exact subpixel stage (agree.hpp:94-185 arithmetic, bit-identical to
// k_agree_subpixel) for the interior matches of this wave's lanes in `pend`
// (row, col, d: each lane's own pixel), run by the whole wave. P =
// min(npts, 64) lanes score one pixel (lane k0 of its group takes the grid
// points k0, k0 + P, ..., x formed by the reference's running float sum),
// so up to min(64 / P, EXACT_GROUPS) pending pixels go per round. Their
// samples come from memory once per round, all loads in flight together,
// into the wave's LDS slice `xs`; both of the reference's passes read them
// there. Each group reduces to the reference's first strict maximum, and
// the owner lane gets the result in res with its ST_PENDING cleared. Must
// be reached by all 64 lanes.
template<typename TIn, int NMAX, int N_EXACT>
__device__ void exact_pending_wave(uint64_t pend, int row, int col, int d, const Stack& f0,
                                   const Stack& f1, Depth<NMAX, N_EXACT> n, float thr, float step,
                                   int npts, bool mv, float minvar, float& res, uint8_t& st,
                                   float4* __restrict__ xs) {
    const int lane = __lane_id();
    const int P = min(npts, 64), G = min(64 / P, EXACT_GROUPS);
    const int g = lane / P, k0 = lane - g * P;
    const int nf = n.get();
    float xl = -1.0f;  // grid point k0
    for (int m = 0; m < k0; ++m) xl += step;
    while (pend) {
        // the next (up to) G pending pixels: group gg takes source lane src[gg]
        int ng = 0, head = -1;
        int prow[EXACT_GROUPS], pcol[EXACT_GROUPS], pc1[EXACT_GROUPS];
#pragma unroll
        for (int gg = 0; gg < EXACT_GROUPS; ++gg) {
            prow[gg] = pcol[gg] = 0;
            pc1[gg] = 1;
            if (gg < G && pend) {
                const int sl = __builtin_ctzll(pend);
                pend &= pend - 1;
                ++ng;
                if (lane == sl) head = gg * P;
                prow[gg] = __builtin_amdgcn_readlane(row, sl);
                pcol[gg] = __builtin_amdgcn_readlane(col, sl);
                pc1[gg] = pcol[gg] - __builtin_amdgcn_readlane(d, sl);
            }
        }
        // stage (u, a, b, c) of every frame of the ng pixels
        constexpr int PER_LANE = (EXACT_GROUPS * NMAX + 63) / 64;
#pragma unroll
        for (int q = 0; q < PER_LANE; ++q) {
            const int e = q * 64 + lane;
            const int gg = e / nf, t = e - gg * nf;
            if (gg < ng) {
                int r = prow[0], c = pcol[0], c1 = pc1[0];
#pragma unroll
                for (int h = 1; h < EXACT_GROUPS; ++h)
                    if (gg == h) {
                        r = prow[h];
                        c = pcol[h];
                        c1 = pc1[h];
                    }
                const float u = (float)load_sample_cached<TIn>(f0, t, r, c);
                const float y0 = (float)load_sample_cached<TIn>(f1, t, r, c1 - 1);
                const float y1 = (float)load_sample_cached<TIn>(f1, t, r, c1);
                const float y2 = (float)load_sample_cached<TIn>(f1, t, r, c1 + 1);
                xs[gg * NMAX + t] = make_float4(u, 0.5f * (y0 - 2.0f * y1 + y2), 0.5f * (-y0 + y2), y1);
            }
        }
        // the wave's LDS operations complete in order: a fence keeps the
        // compiler from moving the reads above the writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool on = g < ng;
        const float4* px = xs + (on ? g : 0) * NMAX;
        float best = -1.0f;
        int best_k = 0x7FFFFFFF;  // none: x stays 0 (agree.hpp:159)
        float x = xl, var0 = 0.0f;
        for (int k = k0; __ballot(on && k < npts); k += P) {
            const float xe = k < npts ? x : 0.0f;  // lanes past the grid: any finite point
            // frame t's interpolated sample at xe (agree.hpp:164)
            const auto sample = [&](const float4& v) {
                return (float)(TIn)(int)__builtin_rintf(v.y * xe * xe + v.z * xe + v.w);
            };
            float mean0 = 0.0f, mean1 = 0.0f;
#pragma unroll 8
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const float4 v = px[t];
                    mean0 += v.x;
                    mean1 += sample(v);
                }
            mean0 /= (float)nf;
            mean1 /= (float)nf;
            float covar = 0.0f, var1 = 0.0f;
            var0 = 0.0f;
#pragma unroll 8
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const float4 v = px[t];
                    const float c0 = v.x - mean0, d1 = sample(v) - mean1;
                    var0 = __builtin_fmaf(c0, c0, var0);
                    covar = __builtin_fmaf(c0, d1, covar);
                    var1 = __builtin_fmaf(d1, d1, var1);
                }
            const float nxc = (mv && var1 < minvar) ? -1.0f : covar / psqrt(var0 * var1);
            if (k < npts && best < nxc) {  // lane-local first strict maximum
                best = nxc;
                best_k = k;
            }
            for (int m = 0; m < P; ++m) x += step;
        }
        // the next round's staging must not overtake this round's reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (mv && var0 < minvar) {  // no grid point is scored (agree.hpp:158)
            best = -1.0f;
            best_k = 0x7FFFFFFF;
        }
        // segmented reduction towards each group's first lane: larger value,
        // then smaller k
        const int seg_end = (g + 1) * P;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const float ob = __shfl_down(best, o);
            const int ok = __shfl_down(best_k, o);
            if (lane + o < seg_end && (ob > best || (ob == best && ok < best_k))) {
                best = ob;
                best_k = ok;
            }
        }
        const float hb = __shfl(best, head < 0 ? lane : head);
        const int hk = __shfl(best_k, head < 0 ? lane : head);
        if (head >= 0) {
            float bx = 0.0f;
            if (hk != 0x7FFFFFFF) {
                bx = -1.0f;
                for (int m = 0; m < hk; ++m) bx += step;
            }
            res = hb < thr ? NAN : (float)d - bx;
            st &= (uint8_t)~ST_PENDING;
        }
    }
}

// Pixels the exact candidate stage takes per round of one wave (each brings
// its candidate grid points, one lane per point, 64 lanes per round). A wave
// has a few pending pixels with one or two candidates each; 8 keeps the
// staging loop short and a workgroup's LDS near 19 KiB at 33 frames.
template<int NMAX>
struct CandLds {
    static constexpr int CAND_PIXELS = 8;
    float4 xs[CAND_PIXELS * NMAX];  // (left sample, a, b, c) of frame t
    int row[CAND_PIXELS], col[CAND_PIXELS], c1[CAND_PIXELS], pre[CAND_PIXELS];
    double mean0[CAND_PIXELS];      // left mean in the correlation's precision
    uint64_t mask[CAND_PIXELS];
    int start[64];
};

__device__ __forceinline__ void wave_lds_sync() {
    // the wave's LDS operations complete in order: a fence keeps the compiler
    // from moving LDS accesses across this point
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The exact subpixel stage (agree.hpp:94-185 arithmetic, bit-identical to
// k_agree_subpixel) restricted to candidate grid points: the lanes with
// `need` set hold in `cmask` (bit k: grid point k) a set of points that
// contains every point whose exact correlation can be the scan's maximum
// (every other point is strictly below some point of the set, so the
// reference's first strict maximum lies in the set and is found by scanning
// the set in grid order). Lanes are handed out one per (pixel, candidate)
// pair: per round, the pending pixels in lane order up to CAND_PIXELS pixels
// or 64 pairs. A round stages each pixel's (left sample - mean0, a, b, c) for
// all frames in LDS, every pair lane runs both of the reference's passes on
// its point there, and each pixel's lanes reduce to the first strict maximum
// (larger correlation, then smaller k). Sets res as the reference would and
// clears need. Must be reached by all 64 lanes; npts <= 64.
//
// inrange: every frame's parabola of the lane's pixel stays inside
// [2, max(TIn) - 2], so the cast of agree.hpp:164 is the identity; a round
// whose pixels all have it skips the cast.
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__device__ void exact_candidates_wave(bool need, uint64_t cmask, int row, int col, int d,
                                      TPrec mean0_own, bool inrange, const Stack& f0,
                                      const Stack& f1, Depth<NMAX, N_EXACT> n, float thr,
                                      const GridX& gx, bool mv, float minvar, float& res,
                                      CandLds<NMAX>& L) {
    constexpr int CAND_PIXELS = CandLds<NMAX>::CAND_PIXELS;
    const int lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    const int nf = n.get();
    const int cnt = need ? __popcll(cmask) : 0;  // 1..64 where need
    while (__ballot(need)) {
        // exclusive prefix of the pair counts over the pending lanes
        const int c = need ? cnt : 0;
        int incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const int pre = incl - c;
        const int rank = __popcll(__ballot(need) & lt);
        const bool take = need && incl <= 64 && rank < CAND_PIXELS;  // a prefix of the pending lanes
        const uint64_t takers = __ballot(take);
        const int last = 63 - __builtin_clzll(takers);
        const int total = __builtin_amdgcn_readlane(incl, last);
        const int ntake = __popcll(takers);
        L.start[lane] = 0;
        wave_lds_sync();
        if (take) {
            L.row[rank] = row;
            L.col[rank] = col;
            L.c1[rank] = col - d;
            L.pre[rank] = pre;
            L.mean0[rank] = (double)mean0_own;
            L.mask[rank] = cmask;
            L.start[pre] = 1;
        }
        wave_lds_sync();
        constexpr int PER_LANE = (CAND_PIXELS * NMAX + 63) / 64;
#pragma unroll 2
        for (int q = 0; q < PER_LANE; ++q) {
            const int e = q * 64 + lane;
            const int gg = e / nf, t = e - gg * nf;
            if (gg < ntake) {
                const int r = L.row[gg], cc = L.col[gg], c1 = L.c1[gg];
                const float u = (float)load_sample_cached<TIn>(f0, t, r, cc);
                const float y0 = (float)load_sample_cached<TIn>(f1, t, r, c1 - 1);
                const float y1 = (float)load_sample_cached<TIn>(f1, t, r, c1);
                const float y2 = (float)load_sample_cached<TIn>(f1, t, r, c1 + 1);
                L.xs[gg * NMAX + t] = make_float4(u, 0.5f * (y0 - 2.0f * y1 + y2), 0.5f * (-y0 + y2), y1);
            }
        }
        // this lane's pair: pixel p (rank among the takers), its j-th candidate
        const bool on = lane < total;
        const uint64_t starts = __ballot(L.start[lane] != 0);
        wave_lds_sync();
        const int p = on ? __popcll(starts & (lt | (1ull << lane))) - 1 : 0;
        const int p_pre = L.pre[p];
        uint64_t m = L.mask[p];
        const int seg_end = p_pre + __popcll(m);
        for (int i = p_pre; i < lane; ++i) m &= m - 1ull;
        const int k = on ? __builtin_ctzll(m) : 0;
        const float xe = gx.x[k];
        const float4* px = L.xs + p * NMAX;
        const TPrec mean0 = (TPrec)L.mean0[p];
        const bool fast = __ballot(take && !inrange) == 0;  // uniform over the wave
        // frame t's interpolated sample at xe (agree.hpp:164), before the cast
        const auto interp = [&](const float4& v) { return __builtin_rintf(v.y * xe * xe + v.z * xe + v.w); };
        const auto wrap = [](float r) { return (float)(TIn)(int)r; };
        // Both of the reference's passes read frame t's sample. The first
        // forms it (with the cast only where a parabola of the round leaves
        // the range; the choice is uniform over the wave) and, for stacks of
        // at most 36 frames, keeps it for the second as a 16-bit value (the
        // samples are integers in [0, max(TIn)]), two per register. The
        // correlation in TPrec as nxcorr / nxcorrd (agree.cuh:35-97).
        constexpr bool KEEP = NMAX <= 36;
        constexpr int UNR = KEEP ? NMAX : 8;
        uint32_t kept[KEEP ? (NMAX + 1) / 2 : 1] = {};
        TPrec mean1 = 0, covar = 0, var0 = 0, var1 = 0;
        const auto pass1 = [&](auto cast) {
#pragma unroll UNR
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const float smp = cast(interp(px[t]));
                    mean1 += (TPrec)smp;
                    if constexpr (KEEP) kept[t >> 1] |= (uint32_t)smp << (16 * (t & 1));
                }
        };
        if (fast) pass1([](float r) { return r; });
        else pass1(wrap);
        mean1 /= (TPrec)nf;
        const auto pass2 = [&](auto sample) {
#pragma unroll UNR
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const TPrec d0 = (TPrec)px[t].x - mean0, d1 = (TPrec)sample(t) - mean1;
                    var0 = pfma(d0, d0, var0);
                    covar = pfma(d0, d1, covar);
                    var1 = pfma(d1, d1, var1);
                }
        };
        if constexpr (KEEP) pass2([&](int t) { return (float)((kept[t >> 1] >> (16 * (t & 1))) & 0xFFFFu); });
        else if (fast) pass2([&](int t) { return interp(px[t]); });
        else pass2([&](int t) { return wrap(interp(px[t])); });
        TPrec best = (mv && (var0 < (TPrec)minvar || var1 < (TPrec)minvar)) ? (TPrec)-1
                                                                             : covar / psqrt(var0 * var1);
        int best_k = k;
        if (!on || !(best > (TPrec)-1)) {  // does not beat the initial -1 (NaN included)
            best = (TPrec)-1;
            best_k = 0x7FFFFFFF;
        }
        // segmented reduction towards each pixel's first lane
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const TPrec ob = __shfl_down(best, o);
            const int ok = __shfl_down(best_k, o);
            if (lane + o < seg_end && (ob > best || (ob == best && ok < best_k))) {
                best = ob;
                best_k = ok;
            }
        }
        const TPrec hb = __shfl(best, take ? pre : lane);
        const int hk = __shfl(best_k, take ? pre : lane);
        if (take) {
            // the reference keeps x = 0 unless some point beats -1 (agree.hpp:159,168-171)
            const float bx = hk != 0x7FFFFFFF ? gx.x[hk] : 0.0f;
            res = hb < (TPrec)thr ? NAN : (float)d - bx;
            need = false;
        }
        wave_lds_sync();  // the next round's staging must not overtake this round's reads
    }
}

// act: the lane's pixel is real (row and col are valid either way; inactive
// lanes compute but write nothing). Must be reached by all 64 lanes.
//
// BX = true: the bit-exact form (Precision::SINGLE without a corrmap,
// npts <= 64). The moments are exact (FP64 sums of integer products), and a
// settled pixel's candidate points are those whose closed form plus bound
// reaches the largest closed form minus bound; any other point's exact
// correlation is strictly below that point's. A single candidate that clears
// the threshold (and -1) by its bound is the reference's result as is;
// otherwise the exact candidate stage scores the candidates (all points for
// an unsettled pixel). Output equals k_agree_subpixel's bit for bit; no
// status is written. `lds` is the wave's CandLds (BX) or float4 slice.
template<typename TIn, int NMAX, int N_EXACT, bool BX = false, typename TPrec = float>
__device__ __forceinline__ void agree_subpixel_moments_px(
    bool act, int row, int col, const int16_t* __restrict__ disp, int64_t dstep, const Stack& f0,
    const Stack& f1, int n_rt, int W, float thr, float step, int npts, int has_minvar, float minvar,
    float* __restrict__ out, int64_t ostep, uint8_t* __restrict__ status, void* lds,
    const GridX* gx = nullptr) {
    using MT = typename std::conditional<BX, double, float>::type;  // cross-moment sums
    const Depth<NMAX, N_EXACT> n{n_rt};
    const bool mv = has_minvar != 0;
    const int d = disp[(size_t)row * dstep + col];
    const int col1 = col - d;
    const bool valid_d = d != INVALID16 && 0 <= col1 && col1 < W;
    const bool interior = valid_d && col1 != 0 && col1 != W - 1;
    const int c1 = interior ? col1 : 1;  // a valid column for the loads of the other lanes
    float res = NAN;
    uint8_t st = 0;

    // One pass: moments of the left samples (every pixel: its variance
    // decides the fallback; FP64, exact for integer samples) and, for
    // interior matches, of the three right columns, with the reach of the
    // parabolas. The cross moments are FP32 sums of samples shifted by the
    // first frame's (products below 2^32; relative error < 1e-5 of the
    // centred moments, covered by the bound's slack below).
    double su = 0, suu = 0;
    MT s0 = 0, s1 = 0, s2 = 0, su0 = 0, su1 = 0, su2 = 0;
    MT s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0, su_f = 0;
    uint64_t cmask = 0;  // BX: candidate grid points of a pending pixel
    int ylo = 0x7FFFFFFF, yhi = -0x7FFFFFFF;
    // branch-free loop (the right samples of lanes without an interior match
    // are read at a valid column and ignored), so that loads of many frames
    // are in flight at once
    // (BX: unshifted, the FP64 sums of integer products are exact)
    const int u_ref = BX ? 0 : (int)load_sample_cached<TIn>(f0, 0, row, col);
    const int y_ref0 = BX ? 0 : (int)load_sample_cached<TIn>(f1, 0, row, c1 - 1);
    const int y_ref1 = BX ? 0 : (int)load_sample_cached<TIn>(f1, 0, row, c1);
    const int y_ref2 = BX ? 0 : (int)load_sample_cached<TIn>(f1, 0, row, c1 + 1);
    // frames in flight per step; fewer for the exact FP64 sums (registers)
    constexpr int MOM_UNROLL = BX ? 8 : 11;
#pragma unroll MOM_UNROLL
    for (int t = 0; t < NMAX; ++t)
        if (n.has(t)) {
            const int ui = (int)load_sample_cached<TIn>(f0, t, row, col);
            const int i0 = (int)load_sample_cached<TIn>(f1, t, row, c1 - 1);
            const int i1 = (int)load_sample_cached<TIn>(f1, t, row, c1);
            const int i2 = (int)load_sample_cached<TIn>(f1, t, row, c1 + 1);
            const double u = (double)ui;
            su += u;
            suu = __builtin_fma(u, u, suu);
            // reach of the frame's parabola on [-1, 1], in integers: the end
            // values and, where it is convex (concave) with the vertex
            // inside, the vertex, which lies above (below) y1 -/+ |y2 - y0| / 4
            // (rounded up here)
            const int r4 = (int)(__builtin_amdgcn_sad_u16((uint32_t)i2, (uint32_t)i0, 3u) >> 2);
            ylo = min(ylo, min(min(i0, i2), i1 - r4));
            yhi = max(yhi, max(max(i0, i2), i1 + r4));
            const MT us = BX ? (MT)u : (MT)(ui - u_ref), y0 = (MT)(i0 - y_ref0), y1 = (MT)(i1 - y_ref1),
                     y2 = (MT)(i2 - y_ref2);
            if constexpr (!BX) su_f += us;
            s0 += y0; s1 += y1; s2 += y2;
            su0 = fma_t(us, y0, su0); su1 = fma_t(us, y1, su1);
            su2 = fma_t(us, y2, su2);
            s00 = fma_t(y0, y0, s00); s01 = fma_t(y0, y1, s01);
            s02 = fma_t(y0, y2, s02); s11 = fma_t(y1, y1, s11);
            s12 = fma_t(y1, y2, s12); s22 = fma_t(y2, y2, s22);
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
            TPrec mean0 = 0;  // the reference's two-pass variance, in its precision
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) mean0 += (TPrec)load_sample_cached<TIn>(f0, t, row, col);
            mean0 /= (TPrec)n.get();
            TPrec var0 = 0;
            for (int t = 0; t < NMAX; ++t)
                if (n.has(t)) {
                    const TPrec c = (TPrec)load_sample_cached<TIn>(f0, t, row, col) - mean0;
                    var0 = pfma(c, c, var0);
                }
            lowvar = var0 < (TPrec)minvar;
        }
    }
    if (lowvar && thr > -1.0f) st |= ST_HOPELESS;
    if (valid_d) {
        if (!interior) {
            const TPrec corr = nxcorr_stream<TIn, TPrec>(f0, row, col, [&](int t) {
                return (float)load_sample_cached<TIn>(f1, t, row, col1);
            }, n, mv, minvar);
            if (!(corr < thr)) res = (float)d;
        } else if (lowvar) {
            if (!(-1.0f < thr)) res = (float)d;  // no grid point is scored (agree.hpp:158)
        } else {
            // centred moments: FP64 from the FP32 sums; BX: n times each is
            // an exact integer in FP64 (n su0 - su s0, ...), rounded once by
            // the division
            const double su_s = BX ? su : (double)su_f, d0 = s0, d1 = s1, d2 = s2;
            const auto centred = [&](double sxy, double sx, double sy) {
                if constexpr (BX) return __builtin_fma((double)n.get(), sxy, -sx * sy) * inv;
                else return sxy - sx * sy * inv;
            };
            const double u0 = centred(su0, su_s, d0), u1 = centred(su1, su_s, d1), u2 = centred(su2, su_s, d2);
            const double g00 = centred(s00, d0, d0), g01 = centred(s01, d0, d1), g02 = centred(s02, d0, d2);
            const double g11 = centred(s11, d1, d1), g12 = centred(s12, d1, d2), g22 = centred(s22, d2, d2);
            const float ua = (float)(0.5 * (u0 - 2.0 * u1 + u2)), ub = (float)(0.5 * (u2 - u0)), uc = (float)u1;
            const float aa = (float)(0.25 * (g00 + 4.0 * g11 + g22 - 4.0 * g01 + 2.0 * g02 - 4.0 * g12));
            const float ab = (float)(0.25 * (g22 - g00 - 2.0 * g12 + 2.0 * g01));
            const float bb = (float)(0.25 * (g00 - 2.0 * g02 + g22));
            const float ac = (float)(0.5 * (g01 - 2.0 * g11 + g12));
            const float bc = (float)(0.5 * (g12 - g01));
            const float cc = (float)g11;
            // Bound on what the left-out rounding can change. If every
            // frame's parabola stays inside [2, max(TIn) - 2] (the integer
            // reach above; the margin covers the FP32 evaluation), the cast of
            // agree.hpp:164 is the identity and every interpolated sample
            // moves by at most 0.52 (0.5 for the rounding plus the FP32
            // evaluation error of ((a x) x + b x) + c, whose four roundings
            // are each at most half an ulp of 65536, 0.0039):
            // |e| <= delta = 0.52 sqrt(n). Then with s = sqrt(var1), the
            // correlation moves by at most `bound` below, and var1 itself by
            // at most 2 s delta + delta^2 (+1e-4 var1 + 1 for the reference's
            // FP32 two-pass variance).
            constexpr float VMAX = (float)(TIn)~(TIn)0;
            const float delta = 0.52f * __builtin_sqrtf((float)n.get());
            float best_x = 0.0f, best = -1.0f, eps_best = 2.0f;
            float upper = -1.0f;  // max over x of closed form + bound
            float lower = -1.0f;  // max over x of closed form - bound
            // BX: largest upper bounds h1 >= h2 >= h3 (at points i1, i2)
            float h1 = -2.0f, h2 = -2.0f, h3 = -2.0f;
            int i1 = 0, i2 = 0, k = 0;
            bool settled = ylo >= 2 && yhi <= (int)VMAX - 2;
            const float rv0 = __builtin_amdgcn_rsqf((float)var0d);
            // The closed form is evaluated with fused multiply-adds (its FP32
            // error is covered by the 1e-4 above); this file is otherwise
            // compiled without contraction for the reference's arithmetic.
            const float ab2 = 2.0f * ab, bbac = bb + 2.0f * ac, bc2 = 2.0f * bc;
            // FP32 evaluation error of the closed forms on |x| <= 1: the
            // coefficients' FP32 rounding plus the fused Horner steps, each
            // at most half an ulp of a partial sum bounded by the sum of the
            // coefficients' magnitudes, so below (5 + 4) 2^-24 = 5.4e-7 of
            // that sum (1e-6 is used): covariance cerr, var1 verr (absolute).
            // Where var1(x) is small against its coefficients this term
            // dominates.
            const float cerr = 1e-6f * (fabsf(ua) + fabsf(ub) + fabsf(uc));
            const float verr = 1e-6f * (fabsf(aa) + fabsf(ab2) + fabsf(bbac) + fabsf(bc2) + fabsf(cc));
            const float cerr_r0 = cerr * rv0;
            const float dd2 = 2.0f * delta, dsq1 = __builtin_fmaf(delta, delta, 1.0f) + verr;
            const auto cov_at = [&](float x) {
                return __builtin_fmaf(__builtin_fmaf(ua, x, ub), x, uc);
            };
            const auto v1_at = [&](float x) {
                return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(__builtin_fmaf(aa, x, ab2), x, bbac), x, bc2), x, cc);
            };
            // Correlation bound at a point with sqrt(var1) = sv > delta,
            // r1 = 1 / sv and closed-form correlation nxc. Rounding the
            // samples adds a centred e, |e| <= delta, to the centred right
            // vector y (|y| = s); with p, q the unit vectors of the centred
            // left samples and of y, r = <p, q>:
            //   r' - r = (<p - r q, e> - r <q, e> (s - s') / (s + s') - r |e|^2 / (s + s')) / s'
            // and |p - r q| = sqrt(1 - r^2), so
            //   |r' - r| <= (sqrt(1 - r^2) delta + delta^2 / (s - delta)) / (s - delta).
            // sqrt(1 - r^2) is taken from nxc with 5e-4 of room for its own
            // error. On top: the reference's FP32 arithmetic on the rounded
            // samples (BX, exact moments: the covariance and variance chains
            // of n fused steps err by at most n 2^-24 relative each, the
            // correlation by below 1.7e-7 n + 1e-6 in all, 1e-6 + 2.5e-7 n
            // is used; MOMENTS, FP32 moments: 1e-4) and the closed forms'
            // error (cerr / sqrt(var0 var1) + verr / var1).
            const float slack = BX ? 1e-6f + 2.5e-7f * (float)n.get() : 1e-4f;
            const float dsq = delta * delta;
            // the bound with w standing for sqrt(1 - r^2); it grows as sv
            // falls and as r1 and w grow, every FP32 step being monotone in
            // its non-negative operands
            const auto bound_w = [&](float sv, float r1, float w) {
                const float inv = 1.000001f * __builtin_amdgcn_rcpf(sv - delta);
                return __builtin_fmaf(__builtin_fmaf(dsq, inv, w * delta), inv,
                                      __builtin_fmaf(r1, __builtin_fmaf(verr, r1, cerr_r0), slack));
            };
            const auto bound = [&](float sv, float r1, float nxc) {
                // v_sqrt_f32 (1 ulp): w only scales the bound
                return bound_w(sv, r1, 1.000001f * __builtin_amdgcn_sqrtf(fmaxf(__builtin_fmaf(-nxc, nxc, 1.0005f), 0.0f)));
            };
            bool full = true;  // the bound at every grid point (below)
            // x = -1 against a constant column col1 - 1 (below): never the maximum
            const bool dead0 = BX && g00 == 0.0;
            // Two levels: TL_K points of the first level are bounded exactly,
            // three for stacks of at most 36 frames (grids of at most 32
            // points), six for deeper stacks (at most 64 points: on the finer
            // grids these usually take, the curve's top points lie within each
            // other's bounds, and six exact points keep the candidate sets
            // small). Other pixels and grids take the per-point scan below.
            // The choice is made per wave: where one of its pixels already
            // fails a condition known before the scan (the parabolas' reach,
            // a left variance above zero), the whole wave takes the per-point
            // scan, which it would run anyway, instead of running both.
            constexpr int TL_K = NMAX > 36 ? 6 : 3;   // exact points of the first level
            constexpr int TL_IB = NMAX > 36 ? 6 : 5;  // bits of a point's index in its key
            constexpr int TL_IM = (1 << TL_IB) - 1;
            if (BX && npts <= TL_IM + 1 && __ballot(!(settled && var0d > 0.0)) == 0) {
                // First the closed form alone at every point, ranked by
                // integer keys: 2^24 times the value truncated toward zero,
                // its low TL_IB bits replaced by the point; the TL_K + 1
                // largest keys (points j_0 .. j_(TL_K - 1) in the low bits of
                // the first TL_K; v_med3 keeps the order statistics), the
                // smallest var1 and whether every |r| is usable. Every other
                // point has its value below nr, the value of the next key
                // rounded up to the next multiple of 2^TL_IB / 2^24. Then the
                // bound at the TL_K points, and for every other point one
                // bound taken at the smallest var1, which is at least that
                // point's own: its sqrt(var1) is not below the smallest one
                // (0.99999 of it is used, far beyond the rounding of v1 *
                // rsq(v1)), its 1 / sqrt(var1) not above (1.00001 of it), and
                // w taken at nr (below). The candidates are then the TL_K
                // points whose upper bounds reach the largest lower bound
                // among them, and every point if nr's upper bound does. A
                // pixel that misses a condition of the bound (sqrt(var1) > 2
                // delta, var1 clear of min_variance by the margin, |r| <=
                // 1.0002, a left variance above zero, more than TL_K grid
                // points, so that the TL_K points take part) takes the
                // per-point scan below. (The grid points come from the kernel
                // arguments, x_k being the reference's running FP32 sum.)
                constexpr float KS = 16777216.0f;  // 2^24
                constexpr int KEY_NONE = -3 * (1 << 24);
                const float rv0s = rv0 * KS;  // exact: a power of two
                // v_cvt_i32_f32: toward zero, saturating, 0 for NaN (a NaN or
                // saturated value also fails |r| <= 1.0002); v_bfi_b32 puts
                // the point (uniform, a scalar register) in the low TL_IB bits
                const auto key_of = [&](float v, int k) {
                    int key;
                    asm("v_cvt_i32_f32 %0, %1\n\tv_bfi_b32 %0, %3, %2, %0" : "=&v"(key) : "v"(v), "s"(k), "i"(TL_IM));
                    return key;
                };
                int kq[TL_K + 1];
#pragma unroll
                for (int i = 0; i <= TL_K; ++i) kq[i] = KEY_NONE;
                float vmin = __builtin_inff();
                bool usable = true;
                // At x = -1 the reference's interpolated samples are the
                // column col1 - 1 itself (a - b + c = y0 exactly in FP32: a, b
                // are half-integers below 2^17). Where that column is constant
                // (g00 = 0, exact) its variance is 0 and the point never wins
                // (-1 under min_variance, else 0 / 0); it takes no part here.
                // (The closed form's error, proportional to the coefficients,
                // cannot show a variance that small.)
                {
                    const float x = gx->x[0];
                    const float v1 = v1_at(x);
                    const float nxs = cov_at(x) * __builtin_amdgcn_rsqf(v1) * rv0s;
                    if (!dead0) {
                        vmin = v1;
                        usable = fabsf(nxs) <= 1.0002f * KS;
                        kq[0] = key_of(nxs, 0);
                    }
                }
                for (int kk = 1; kk < npts; ++kk) {
                    const float x = gx->x[kk];
                    const float v1 = v1_at(x);
                    const float nxs = cov_at(x) * __builtin_amdgcn_rsqf(v1) * rv0s;  // 2^24 r
                    // smallest var1 as an integer min of the bit patterns: it
                    // orders non-negative values like floats, and a negative
                    // (or NaN) var1 comes out negative or leaves |r| <= 1.0002
                    // false, either of which sends the pixel to the per-point scan
                    vmin = __int_as_float(min(__float_as_int(vmin), __float_as_int(v1)));
                    usable = usable && fabsf(nxs) <= 1.0002f * KS;  // false for NaN
                    const int key = key_of(nxs, kk);
#pragma unroll
                    for (int i = TL_K; i >= 1; --i) kq[i] = imed3(kq[i - 1], kq[i], key);
                    kq[0] = max(kq[0], key);
                }
                int js[TL_K];
#pragma unroll
                for (int i = 0; i < TL_K; ++i) js[i] = kq[i] & TL_IM;
                const float nr = (float)((kq[TL_K] & ~TL_IM) + TL_IM + 1) * (1.0f / KS);  // exact
                const float r1m = __builtin_amdgcn_rsqf(vmin), svm = vmin * r1m;
                // var1 - margin(var1) rises with var1 where sqrt(var1) > delta, so
                // the smallest var1 clearing min_variance by its margin (+1 for
                // the rounding of the margin) clears it at every point
                const bool clear_mv = !mv || vmin >= minvar + __builtin_fmaf(dd2, 1.00001f * svm,
                                                                             __builtin_fmaf(1e-4f, vmin, dsq1)) + 1.0f;
                if (settled && usable && npts >= TL_K + 1 && var0d > 0.0 && vmin > 4.0004f * dsq && clear_mv) {
                    full = false;
                    const uint64_t all = npts >= 64 ? ~0ull : (1ull << npts) - 1ull;
                    const auto at = [&](int j, float& hi) {  // lower bound at point j, its upper in hi
                        const float x = gx->x[j];
                        const float v1 = v1_at(x), r1 = __builtin_amdgcn_rsqf(v1);
                        const float nxc = cov_at(x) * r1 * rv0;
                        const float e = bound(v1 * r1, r1, nxc);
                        hi = nxc + e;
                        return nxc - e;
                    };
                    float his[TL_K];
                    float lower = -__builtin_inff();
#pragma unroll
                    for (int i = 0; i < TL_K; ++i) lower = fmaxf(lower, at(js[i], his[i]));
                    // Every other point has its value v <= nr, and its upper
                    // bound is below v + a + c w(v) with c = delta inv, a the
                    // rest of the bound and inv, r1 taken at the smallest var1.
                    // v + c w(v) rises with v wherever c v < sqrt(1.0005 - v^2),
                    // so if that holds at nr (with room), w(nr) serves for all
                    // of them (1.0002 covers the FP32 square roots); else w <=
                    // 1.0003.
                    const float svl = 0.99999f * svm;
                    const float cw = 1.000001f * delta * (1.000001f * __builtin_amdgcn_rcpf(svl - delta));
                    const float sqr = __builtin_amdgcn_sqrtf(fmaxf(__builtin_fmaf(-nr, nr, 1.0005f), 0.0f));
                    const float wr = cw * nr <= 0.999f * sqr ? 1.0002f * sqr : 1.0003f;
                    const float hir = nr + bound_w(svl, 1.00001f * r1m, wr);
                    float hmax = fmaxf(hir, -1.0f);  // the reference's best starts at -1
#pragma unroll
                    for (int i = 0; i < TL_K; ++i) hmax = fmaxf(hmax, his[i]);
                    if (!(hmax < thr)) {
                        if (hir >= lower) {
                            cmask = all;
                        } else {
                            uint64_t m = 1ull << js[0];
                            bool more = false;
#pragma unroll
                            for (int i = 1; i < TL_K; ++i)
                                if (his[i] >= lower) {
                                    m |= 1ull << js[i];
                                    more = true;
                                }
                            // one candidate: lower is j_0's lower bound (the point
                            // attaining it is a candidate), its exact value is at
                            // least that and above every other point's
                            if (!more && lower >= thr && lower > -1.0f) res = (float)d - gx->x[js[0]];
                            else cmask = m;
                        }
                        if (cmask) st |= ST_PENDING;
                    }
                }
            }
            if (full) {
            for (float x = -1.0f; x <= 1.0f; x += step) {
                const float cov = cov_at(x);
                const float v1 = v1_at(x);
                const float r1 = __builtin_amdgcn_rsqf(v1);
                const float sv = v1 > 0.0f ? v1 * r1 : 0.0f;  // sqrt(var1)
                const float margin = __builtin_fmaf(dd2, sv, __builtin_fmaf(1e-4f, fabsf(v1), dsq1));
                float nxc = -1.0f, eps = 0.0f;
                if (dead0 && k == 0) {
                    // x = -1 against a constant column: -1 or NaN, never the maximum
                } else if (mv && v1 < minvar) {
                    // below the variance test: also exactly when below by the margin
                    if (v1 >= minvar - margin) settled = false;
                } else {
                    if (mv && v1 < minvar + margin) settled = false;
                    if (!(sv > 2.0f * delta)) settled = false;  // no useful bound here
                    nxc = cov * r1 * rv0;
                    eps = bound(sv, r1, nxc);
                }
                if (!(fabsf(nxc) <= 1.0002f)) settled = false;  // no usable closed form
                const float hi = nxc + eps;
                lower = fmaxf(lower, nxc - eps);
                if constexpr (BX) {
                    // the three largest upper bounds, with the points of the
                    // first two (v_med3 keeps the order statistics)
                    const bool g1 = hi > h1, g2 = hi > h2;
                    i2 = g1 ? i1 : (g2 ? k : i2);
                    i1 = g1 ? k : i1;
                    h3 = fmaxf(fminf(hi, h2), h3);
                    h2 = fmaxf(fminf(hi, h1), h2);
                    h1 = fmaxf(h1, hi);
                    ++k;
                } else {
                    upper = fmaxf(upper, hi);
                    if (best < nxc) {
                        best = nxc;
                        best_x = x;
                        eps_best = eps;
                    }
                }
            }
            if constexpr (BX) {
                // Candidates are the points whose upper bound reaches `lower`
                // (the point attaining `lower` is one of them). Only the first
                // (h2 < lower) or the first two (h3 < lower) upper bounds can
                // qualify; a wider set is scored in full.
                const uint64_t all = npts >= 64 ? ~0ull : (1ull << npts) - 1ull;
                if (!settled) {
                    cmask = all;
                    st |= ST_PENDING;
                } else if (!(fmaxf(h1, -1.0f) < thr)) {  // the reference's best starts at -1
                    if (h2 < lower) {
                        // one candidate: its exact value is at least `lower`
                        // and above every other point's
                        if (lower >= thr && lower > -1.0f) res = (float)d - gx->x[i1];
                        else cmask = 1ull << i1;
                    } else {
                        cmask = h3 < lower ? (1ull << i1) | (1ull << i2) : all;
                    }
                    if (cmask) st |= ST_PENDING;
                }
            } else if (settled && best - eps_best >= thr) {
            // Accepted with the closed-form grid point when the bound settles
            // it and no grid point half a pixel or more away could beat the
            // winner; else the exact stage.
                float far = -1.0f;
                for (float x = -1.0f; x <= 1.0f; x += step) {
                    if (fabsf(x - best_x) < 0.5f) continue;
                    const float v1 = v1_at(x);
                    const float r1 = __builtin_amdgcn_rsqf(v1);
                    if (mv && v1 < minvar) continue;  // settled: exactly -1 there
                    const float nxc = cov_at(x) * r1 * rv0;
                    far = fmaxf(far, nxc + bound(v1 * r1, r1, nxc));
                }
                if (far < best - eps_best)
                    res = (float)d - best_x;
                else
                    st |= ST_PENDING;
            } else if (!(settled && upper < thr)) {
                st |= ST_PENDING;
            }
            }  // full
        }
    }
    const uint64_t pend = __ballot(act && (st & ST_PENDING));
    if constexpr (BX) {
        if (pend) {
            // the reference's left mean in TPrec: su is the exact sample sum
            const TPrec mean0 = (TPrec)su / (TPrec)n.get();
            const bool inrange = ylo >= 2 && yhi <= (int)(TIn)~(TIn)0 - 2;
            exact_candidates_wave<TIn, TPrec>(act && (st & ST_PENDING), cmask, row, col, d, mean0,
                                              inrange, f0, f1, n, thr, *gx, mv, minvar, res,
                                              *(CandLds<NMAX>*)lds);
        }
        if (act) out[(size_t)row * ostep + col] = res;
    } else {
        if (pend)
            exact_pending_wave<TIn>(pend, row, col, d, f0, f1, n, thr, step, npts, mv, minvar, res, st,
                                    (float4*)lds);
        if (act) {
            out[(size_t)row * ostep + col] = res;
            status[(size_t)row * W + col] = st;
        }
    }
}

template<typename TIn, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_moments(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int npts, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, uint8_t* __restrict__ status) {
    __shared__ float4 xs[4][EXACT_GROUPS * NMAX];
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if ((col & ~63) >= W)  // whole wave past the row
        return;
    agree_subpixel_moments_px<TIn, NMAX, N_EXACT>(col < W, row, min(col, W - 1), disp, dstep, f0, f1,
                                                  n_rt, W, thr, step, npts, has_minvar, minvar, out,
                                                  ostep, status, xs[threadIdx.x >> 6]);
}

// Bit-exact subpixel agreement (Precision::SINGLE with TPrec = float,
// Precision::DOUBLE with TPrec = double; no corrmap, npts <= 64):
// closed-form pruning, the exact stage on the candidate points only.
template<typename TIn, typename TPrec, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_pruned(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int npts, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, const GridX gx) {
    __shared__ CandLds<NMAX> cl[4];
    // the workgroups of one row on consecutive slots of one XCD: its L2 then
    // holds the right-view lines that neighbouring workgroups share
    int row, bx;
    xcd_row_group(row, bx);
    const int col = bx * 256 + threadIdx.x;
    if ((col & ~63) >= W)  // whole wave past the row
        return;
    agree_subpixel_moments_px<TIn, NMAX, N_EXACT, true, TPrec>(col < W, row, min(col, W - 1), disp, dstep,
                                                               f0, f1, n_rt, W, thr, step, npts, has_minvar,
                                                               minvar, out, ostep, nullptr,
                                                               &cl[threadIdx.x >> 6], &gx);
}

// over listed pixels only (dense numbering of the lists; each wave walks
// its 64 entries per round, lanes past the end repeat the last entry)
template<typename TIn, int NMAX, int N_EXACT>
__global__ __launch_bounds__(256) void k_agree_subpixel_moments_list(
    const int16_t* __restrict__ disp, int64_t dstep, const Stack f0, const Stack f1, int n_rt,
    int W, float thr, float step, int npts, int has_minvar, float minvar, float* __restrict__ out,
    int64_t ostep, uint8_t* __restrict__ status, const RowLists l) {
    __shared__ float4 xs[4][EXACT_GROUPS * NMAX];
    const int total = l.row_start[l.H];
    const int lane = threadIdx.x & 63;
    for (int wbase = blockIdx.x * 256 + (threadIdx.x & ~63); wbase < total; wbase += gridDim.x * 256) {
        const int i = wbase + lane;
        int row, col;
        list_entry(l, min(i, total - 1), row, col);
        // LISTED_SAME16: keeps the windowed pass's outputs (inactive lane)
        const bool act = i < total && disp[(size_t)row * dstep + col] != LISTED_SAME16;
        agree_subpixel_moments_px<TIn, NMAX, N_EXACT>(act, row, col, disp, dstep, f0, f1, n_rt, W,
                                                      thr, step, npts, has_minvar, minvar, out, ostep,
                                                      status, xs[threadIdx.x >> 6]);
    }
}

}  // namespace bicos_hip
