// Windowed Hamming search and full-search fallback of the learned-window
// matching path (no reference counterpart; the search semantics restricted
// to a window are those of include/impl/cpu/bicos.hpp:50-113).
//
// Every pixel searches the candidate columns [c - hw, c + hw] (clipped to the
// row) around c = col - round(centre) of its 4 x 4 cell's window. Keys and
// the duplicate rule are the full search's (search.hpp), restricted to the
// window. Pixels whose windowed match is then rejected by the NCC stage are
// listed (k_fallback_list) and searched over the whole row
// (k_search_list), and their NCC stage is run again on the list.
#pragma once

#include "agree.hpp"
#include "agree_fast.hpp"
#include "agree_moments.hpp"
#include "cnn.hpp"
#include "device_common.hpp"
#include "search.hpp"
#include "search_mfma.hpp"

namespace bicos_hip {

template<int WORDS>
__device__ __forceinline__ void load_desc(uint32_t (&q)[WORDS], const uint32_t* p) {
    if constexpr (WORDS >= 4) {
#pragma unroll
        for (int k = 0; k < WORDS; k += 4) {
            const uint4 v = *(const uint4*)(p + k);
            q[k] = v.x; q[k + 1] = v.y; q[k + 2] = v.z; q[k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < WORDS; ++k) q[k] = p[k];
    }
}

// two smallest keys of the candidates [lo, hi] of row `crow` against q
template<int WORDS, bool NODUPES>
__device__ __forceinline__ void scan_window(const uint32_t (&q)[WORDS], const uint32_t* crow, int lo,
                                            int hi, uint32_t& b1, uint32_t& b2) {
    b1 = 0xFFFFFFFFu;
    b2 = 0xFFFFFFFFu;
    for (int c = lo; c <= hi; ++c) {
        uint32_t d[WORDS];
        load_desc<WORDS>(d, crow + (size_t)c * WORDS);
        uint32_t cost = 0;
#pragma unroll
        for (int k = 0; k < WORDS; ++k) cost = __popc(q[k] ^ d[k]) + cost;
        const uint32_t key = (cost << 16) | (uint32_t)c;
        if constexpr (NODUPES) b2 = umed3(b1, b2, key);
        b1 = min(b1, key);
    }
}

// int16 disparity from the windowed search. CONSIST: the matched right
// column searches the left row back over [col - hw, col + hw] and the match
// stands if that reverse best lies within max_lr_diff of col
// (bicos.hpp:96-106 restricted to the window).
template<int WORDS, bool NODUPES, bool CONSIST>
__global__ __launch_bounds__(256) void k_window_search(const uint32_t* __restrict__ d0,
                                                       const uint32_t* __restrict__ d1, int W,
                                                       const float* __restrict__ center_q,
                                                       const int16_t* __restrict__ hw_q, int Hq,
                                                       int Wq, int max_lr_diff,
                                                       int16_t* __restrict__ out, int64_t ostep) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;
    const int cq = min(row / CNN_DS, Hq - 1) * Wq + min(col / CNN_DS, Wq - 1);
    const int c1 = col - (int)__builtin_rintf(center_q[cq]);
    const int hw = hw_q[cq];
    const int lo = max(c1 - hw, 0), hi = min(c1 + hw, W - 1);
    uint32_t q[WORDS];
    load_desc<WORDS>(q, d0 + ((size_t)row * W + col) * WORDS);
    uint32_t b1, b2;
    scan_window<WORDS, NODUPES>(q, d1 + (size_t)row * W * WORDS, lo, hi, b1, b2);
    int16_t v = INVALID16;
    const bool valid = lo <= hi && !(NODUPES && (b1 >> 16) == (b2 >> 16));
    if (valid) {
        const int best = (int)(b1 & 0xFFFFu);
        if constexpr (!CONSIST) {
            v = (int16_t)(col - best);
        } else {
            uint32_t r[WORDS];
            load_desc<WORDS>(r, d1 + ((size_t)row * W + best) * WORDS);
            uint32_t r1, r2;
            scan_window<WORDS, NODUPES>(r, d0 + (size_t)row * W * WORDS, max(col - hw, 0),
                                        min(col + hw, W - 1), r1, r2);
            const int r0 = (int)(r1 & 0xFFFFu);
            if (!(NODUPES && (r1 >> 16) == (r2 >> 16)) && abs(col - r0) <= max_lr_diff)
                v = (int16_t)((col + r0) / 2 - best);
        }
    }
    out[(size_t)row * ostep + col] = v;
}

// Two smallest keys of the candidates [lo, hi] of one row, read from `src`
// (LDS or global memory) at column c - base; columns are 16-byte aligned
// for WORDS >= 4 (read as uint4).
template<int WORDS, bool NODUPES>
__device__ __forceinline__ void scan_keys(const uint32_t (&q)[WORDS], const uint32_t* src, int base,
                                          int lo, int hi, uint32_t& b1, uint32_t& b2) {
    for (int c = lo; c <= hi; ++c) {
        uint32_t d[WORDS];
        if constexpr (WORDS >= 4) {
            const uint4* p4 = (const uint4*)src + (size_t)(c - base) * (WORDS / 4);
#pragma unroll
            for (int k = 0; k < WORDS / 4; ++k) {
                const uint4 v = p4[k];
                d[4 * k] = v.x; d[4 * k + 1] = v.y; d[4 * k + 2] = v.z; d[4 * k + 3] = v.w;
            }
        } else {
            load_desc<WORDS>(d, src + (size_t)(c - base) * WORDS);
        }
        uint32_t cost = 0;
#pragma unroll
        for (int k = 0; k < WORDS; ++k) cost = __popc(q[k] ^ d[k]) + cost;
        const uint32_t key = (cost << 16) | (uint32_t)c;
        if constexpr (NODUPES) b2 = umed3(b1, b2, key);
        b1 = min(b1, key);
    }
}

// The window scan of one pixel over [lo, hi]: columns [base, base + len) of
// the row are staged in `lds`, the rest is read from `grow` (the row in
// global memory). Candidates are taken in column order (the first strict
// minimum needs it): the staged part, then the columns after it.
template<int WORDS, bool NODUPES>
__device__ __forceinline__ void scan_staged(const uint32_t (&q)[WORDS], const uint32_t* lds, int base,
                                            int len, const uint32_t* grow, int lo, int hi,
                                            uint32_t& b1, uint32_t& b2) {
    b1 = 0xFFFFFFFFu;
    b2 = 0xFFFFFFFFu;
    const int split = base + len;  // lo >= base always
    if (lo < base) {
        scan_keys<WORDS, NODUPES>(q, grow, 0, lo, hi, b1, b2);
        return;
    }
    scan_keys<WORDS, NODUPES>(q, lds, base, lo, min(hi, split - 1), b1, b2);
    scan_keys<WORDS, NODUPES>(q, grow, 0, max(lo, split), hi, b1, b2);
}

// Copies `total` 16-byte units from global memory to LDS, ld(u) -> uint4
// and st(u, v), B units per thread in flight at once.
template<int B, typename Ld, typename St>
__device__ __forceinline__ void stage_units(int total, Ld&& ld, St&& st) {
    for (int u0 = (int)threadIdx.x; u0 < total; u0 += 256 * B) {
        uint4 v[B];
#pragma unroll
        for (int b = 0; b < B; ++b)
            if (u0 + 256 * b < total) v[b] = ld(u0 + 256 * b);
#pragma unroll
        for (int b = 0; b < B; ++b)
            if (u0 + 256 * b < total) st(u0 + 256 * b, v[b]);
    }
}

// LDS of k_window_core for a span of `span` staged columns (a multiple of 8)
__host__ __device__ constexpr size_t core_lds_bytes(int words, bool consist, int span) {
    return (size_t)span * words * 4 * (consist ? 2 : 1);
}

// The learned path's windowed pass in one kernel. A block takes 256 pixels
// of one row and stages in LDS, all loads in flight together, the right
// descriptors of the columns its windows reach (up to `span` columns from
// the first window's start, less one, rounded down to 8; windows reaching
// past them read the rest from global memory) and, with Consistency, the
// left descriptors of its back-search windows. Then per pixel: the windowed
// search (k_window_search's semantics), the MOMENTS_FAST subpixel stage on
// its match, and the fallback set (invalid result; not a pixel whose left
// variance rejects every match; with `unsure`, only cells given the wide
// window) appended to the row's list (row_count[row] entries at
// list + row * W, zeroed beforehand), its windowed match kept in dwin for
// the full-row search of the list. npts <= 64 grid points.
template<typename TIn, int WORDS, bool NODUPES, bool CONSIST>
__global__ __launch_bounds__(256) void k_window_core(
    const uint32_t* __restrict__ d0, const uint32_t* __restrict__ d1, const float* __restrict__ center_q,
    const int16_t* __restrict__ hw_q, int Hq, int Wq, int max_lr_diff, const Stack f0, const Stack f1,
    int n, int W, float thr, int npts, const GridX gx, int has_mv, float minvar,
    float* __restrict__ out, int64_t ostep, int fallback, int unsure, int hw_lo,
    int16_t* __restrict__ list, int* __restrict__ row_count, int span, int16_t* __restrict__ dwin) {
    extern __shared__ uint4 core_lds[];
    uint32_t* rdesc = (uint32_t*)core_lds;                    // span x WORDS
    uint32_t* ldesc = rdesc + (size_t)span * WORDS;           // span x WORDS (Consistency)
    __shared__ int red[4][3], wsum[4], base_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int row, bx;
    xcd_row_group_il(row, bx);  // one row's workgroups on one XCD, rows dealt out over the XCDs
    const int c0 = bx * 256;
    const int col = min(c0 + tid, W - 1);
    const bool act = c0 + tid < W;

    const int cq = min(row / CNN_DS, Hq - 1) * Wq + min(col / CNN_DS, Wq - 1);
    const int k1 = col - (int)__builtin_rintf(center_q[cq]);
    const int hw = hw_q[cq];
    const int lo = max(k1 - hw, 0), hi = min(k1 + hw, W - 1);
    // block extent of the windows and the largest half-width
    int mlo = act && lo <= hi ? lo : 0x7FFFFFFF, mhi = act && lo <= hi ? hi : -1;
    int mhw = act ? hw : 0;
    for (int o = 32; o > 0; o >>= 1) {
        mlo = min(mlo, __shfl_xor(mlo, o));
        mhi = max(mhi, __shfl_xor(mhi, o));
        mhw = max(mhw, __shfl_xor(mhw, o));
    }
    if (lane == 0) {
        red[wave][0] = mlo;
        red[wave][1] = mhi;
        red[wave][2] = mhw;
    }
    __syncthreads();
    const int rbase = min(min(red[0][0], red[1][0]), min(red[2][0], red[3][0]));
    const int rlast = max(max(red[0][1], red[1][1]), max(red[2][1], red[3][1]));
    // staged right columns [sa, sa + slen): the windows' reach plus one column
    // each side (the subpixel stage reads col1 - 1 .. col1 + 1)
    const int sa = rlast >= 0 ? max(rbase - 1, 0) & ~7 : 0;
    const int slen = rlast >= 0 ? min(min(rlast + 1, W - 1) - sa + 1, span) : 0;
    const uint32_t* rrow = d1 + (size_t)row * W * WORDS;
    const uint32_t* lrow = d0 + (size_t)row * W * WORDS;
    const auto stage_desc = [&](const uint32_t* src, int first, int len, uint32_t* dst) {
        if constexpr (WORDS >= 4) {
            stage_units<4>(len * (WORDS / 4), [&](int u) { return ((const uint4*)(src + (size_t)first * WORDS))[u]; },
                           [&](int u, uint4 v) { ((uint4*)dst)[u] = v; });
        } else {
            for (int i = tid; i < len * WORDS; i += 256) dst[i] = src[(size_t)first * WORDS + i];
        }
    };
    stage_desc(rrow, sa, slen, rdesc);
    int lb = 0, llen = 0;
    if constexpr (CONSIST) {
        const int bhw = max(max(red[0][2], red[1][2]), max(red[2][2], red[3][2]));
        lb = max(c0 - bhw, 0);
        llen = min(min(c0 + 255 + bhw, W - 1) - lb + 1, span);
        stage_desc(lrow, lb, llen, ldesc);
    }
    __syncthreads();

    uint32_t q[WORDS];
    load_desc<WORDS>(q, lrow + (size_t)col * WORDS);
    uint32_t b1, b2;
    scan_staged<WORDS, NODUPES>(q, rdesc, sa, slen, rrow, lo, hi, b1, b2);
    int d = INVALID16;
    if (lo <= hi && !(NODUPES && (b1 >> 16) == (b2 >> 16))) {
        const int best = (int)(b1 & 0xFFFFu);
        if constexpr (!CONSIST) {
            d = col - best;
        } else {
            uint32_t r[WORDS];
            load_desc<WORDS>(r, rrow + (size_t)best * WORDS);
            uint32_t r1, r2;
            const int blo = max(col - hw, 0), bhi = min(col + hw, W - 1);
            if (blo >= lb)
                scan_staged<WORDS, NODUPES>(r, ldesc, lb, llen, lrow, blo, bhi, r1, r2);
            else
                scan_staged<WORDS, NODUPES>(r, ldesc, blo, 0, lrow, blo, bhi, r1, r2);
            const int r0 = (int)(r1 & 0xFFFFu);
            if (!(NODUPES && (r1 >> 16) == (r2 >> 16)) && abs(col - r0) <= max_lr_diff)
                d = (col + r0) / 2 - best;
        }
    }

    const FastCols c(col, d, W);
    FastMoments m;
#pragma unroll 8
    for (int t = 0; t < n; ++t)
        m.add((double)load_sample_cached<TIn>(f0, t, row, col), (double)load_sample_cached<TIn>(f1, t, row, c.cl),
              (double)load_sample_cached<TIn>(f1, t, row, c.cmid), (double)load_sample_cached<TIn>(f1, t, row, c.cr));
    bool lowvar;
    const float res = subpixel_fast_decide(m, n, d, c.valid_d, c.interior, thr, npts, gx, has_mv != 0,
                                           minvar, lowvar);
    if (act) out[(size_t)row * ostep + col] = res;
    if (!fallback) return;  // uniform
    if (act) dwin[(size_t)row * W + col] = (int16_t)d;  // the windowed match, for set_listed
    const bool take = act && __builtin_isnan(res) && !(lowvar && thr > -1.0f) && !(unsure && hw == hw_lo);
    const uint64_t mk = __ballot(take);
    if (lane == 0) wsum[wave] = __popcll(mk);
    __syncthreads();
    if (tid == 0) {
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        base_s = total ? atomicAdd(&row_count[row], total) : 0;
    }
    __syncthreads();
    if (take) {
        int off = base_s + __popcll(mk & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) off += wsum[w];
        list[(size_t)row * W + off] = (int16_t)col;
    }
}

// Whether pixel (row, col) lies in a cell the network was unsure about: it
// got the wide window (half-width other than hw_lo)
__device__ __forceinline__ bool cell_unsure(const int16_t* hw_q, int hw_lo, int Hq, int Wq, int row,
                                            int col) {
    return hw_q[min(row / CNN_DS, Hq - 1) * Wq + min(col / CNN_DS, Wq - 1)] != hw_lo;
}

// Fallback set: pixels whose windowed result is invalid and whose left
// variance does not already reject every match (has_minvar and
// var0 < minvar, computed as the NCC stage does in precision TPrec), and,
// with hw_q given, that lie in a cell the network was unsure about. Each
// block takes 256 columns of one row (the row is blockIdx.x so that rows
// spread over the XCDs) and appends its pixels to the row's list
// (row_count[row] entries at list + row * W, in no particular order);
// row_count starts at zero.
template<typename TIn, typename TPrec>
__global__ __launch_bounds__(256) void k_fallback_list(const float* __restrict__ disp_f,
                                                       const int16_t* __restrict__ disp_i,
                                                       int64_t dstep, const Stack f0, int n, int W,
                                                       int has_minvar, float minvar, float thr,
                                                       const int16_t* __restrict__ hw_q, int hw_lo,
                                                       int Hq, int Wq, int16_t* __restrict__ list,
                                                       int* __restrict__ row_count) {
    __shared__ int wsum[4], base;
    const int row = blockIdx.x, col = blockIdx.y * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool need = false;
    if (col < W) {
        const size_t i = (size_t)row * dstep + col;
        need = disp_f ? __builtin_isnan(disp_f[i]) : disp_i[i] == INVALID16;
        if (need && hw_q) need = cell_unsure(hw_q, hw_lo, Hq, Wq, row, col);
        if (need && has_minvar && thr > -1.0f) {
            // var0 as the NCC stage forms it: mean, then an fma chain of the
            // centred samples in frame order
            float v[MAX_FRAMES];
#pragma unroll
            for (int t = 0; t < MAX_FRAMES; ++t)
                v[t] = t < n ? (float)load_sample_cached<TIn>(f0, t, row, col) : 0.0f;
            TPrec mean = 0;
#pragma unroll
            for (int t = 0; t < MAX_FRAMES; ++t)
                if (t < n) mean += (TPrec)v[t];
            mean /= (TPrec)n;
            TPrec var = 0;
#pragma unroll
            for (int t = 0; t < MAX_FRAMES; ++t)
                if (t < n) {
                    const TPrec d = (TPrec)v[t] - mean;
                    var = pfma(d, d, var);
                }
            need = !(var < (TPrec)minvar);
        }
    }
    const uint64_t m = __ballot(need);
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        base = total ? atomicAdd(&row_count[row], total) : 0;
    }
    __syncthreads();
    if (need) {
        int off = __popcll(m & ((1ull << lane) - 1));
        for (int w = 0; w < wave; ++w) off += wsum[w];
        list[(size_t)row * W + base + off] = (int16_t)col;
    }
}

// Fallback set from the MOMENTS status: invalid result and not hopeless,
// and, with hw_q given, in a cell the network was unsure about. As
// k_fallback_list without touching the stacks.
__global__ __launch_bounds__(256) void k_fallback_list_status(
    const uint8_t* __restrict__ status, const float* __restrict__ res, int64_t rstep, int W,
    const int16_t* __restrict__ hw_q, int hw_lo, int Hq, int Wq, int16_t* __restrict__ list,
    int* __restrict__ row_count) {
    __shared__ int wsum[4], base;
    const int row = blockIdx.x, col = blockIdx.y * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool take = false;
    if (col < W) {
        take = __builtin_isnan(res[(size_t)row * rstep + col]) && !(status[(size_t)row * W + col] & ST_HOPELESS);
        if (take && hw_q)
            take = cell_unsure(hw_q, hw_lo, Hq, Wq, row, col);
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

// Full-row search of the listed query columns of each row (the row's
// candidates staged in LDS as in k_search). Entry i of row r queries column
// qcol[r * W + i] (skipped when negative) and writes
//   MODE SEARCH_REV:  the best candidate column (INVALID16 if none/duplicate)
//                     to res[r * W + i]
//   MODE SEARCH_DISP: col - best to disp[r * dstep + col]
// One block serves 256 entries of one row; waves past the row's count skip
// the arithmetic but take part in the staging. The row is blockIdx.x:
// consecutive workgroups go to different XCDs, and most rows need only
// their first block, so rows (not entry blocks) must be what the
// round-robin spreads.
template<int WORDS, bool NODUPES, int MODE>
__global__ __launch_bounds__(SEARCH_THREADS) void k_search_list(
    const uint32_t* __restrict__ dq, const uint32_t* __restrict__ dc, int W, int chunk_cols,
    const int16_t* __restrict__ qcol, const int* __restrict__ row_count,
    int16_t* __restrict__ res, int16_t* __restrict__ disp, int64_t dstep) {
    extern __shared__ uint4 lds4[];
    uint32_t* lds = (uint32_t*)lds4;
    const int row = blockIdx.x, tid = threadIdx.x;
    const int count = row_count[row];
    const int i0 = blockIdx.y * SEARCH_THREADS;
    if (i0 >= count)
        return;  // uniform over the block
    const int i = i0 + tid;
    const int col = i < count ? qcol[(size_t)row * W + i] : -1;
    const bool wave_busy = i0 + (tid & ~63) < count;
    uint32_t q[WORDS];
    load_desc<WORDS>(q, dq + ((size_t)row * W + (col >= 0 ? col : 0)) * WORDS);
    uint32_t b1 = 0xFFFFFFFFu, b2 = 0xFFFFFFFFu;
    const uint32_t* crow = dc + (size_t)row * W * WORDS;
    for (int c0 = 0; c0 < W; c0 += chunk_cols) {
        const int cn = min(chunk_cols, W - c0);
        __syncthreads();
        if constexpr (WORDS >= 4) {
            const uint4* src = (const uint4*)(crow + (size_t)c0 * WORDS);
            for (int k = tid; k < cn * (WORDS / 4); k += SEARCH_THREADS) lds4[k] = src[k];
        } else {
            const uint32_t* src = crow + (size_t)c0 * WORDS;
            for (int k = tid; k < cn * WORDS; k += SEARCH_THREADS) lds[k] = src[k];
        }
        __syncthreads();
        if (wave_busy) {
#pragma unroll 4
            for (int c = 0; c < cn; ++c) {
                uint32_t cost = 0;
#pragma unroll
                for (int k = 0; k < WORDS; ++k) cost = __popc(q[k] ^ lds[c * WORDS + k]) + cost;
                const uint32_t key = (cost << 16) | (uint32_t)(c0 + c);
                if constexpr (NODUPES) b2 = umed3(b1, b2, key);
                b1 = min(b1, key);
            }
        }
    }
    if (i >= count)
        return;
    const bool valid = col >= 0 && !(NODUPES && (b1 >> 16) == (b2 >> 16));
    const int best = (int)(b1 & 0xFFFFu);
    if constexpr (MODE == SEARCH_REV) {
        res[(size_t)row * W + i] = valid ? (int16_t)best : INVALID16;
    } else {
        if (col >= 0) set_listed(disp + (size_t)row * dstep + col, valid ? (int16_t)(col - best) : INVALID16);
    }
}

// Consistency fallback: entry i of row r matched left column list[i] with
// right column fwd[i]; rev[i] is that right column's best left column.
__global__ __launch_bounds__(256) void k_lr_combine(const int16_t* __restrict__ list,
                                                    const int* __restrict__ row_count, int W,
                                                    const int16_t* __restrict__ fwd,
                                                    const int16_t* __restrict__ rev, int max_lr_diff,
                                                    int16_t* __restrict__ disp, int64_t dstep) {
    const int row = blockIdx.x, count = row_count[row];  // rows spread over XCDs
    for (int i = blockIdx.y * 256 + threadIdx.x; i < count; i += gridDim.y * 256) {
        const size_t e = (size_t)row * W + i;
        const int col = list[e], best = fwd[e], r0 = rev[e];
        int16_t v = INVALID16;
        if (best != INVALID16 && r0 != INVALID16 && abs(col - r0) <= max_lr_diff)
            v = (int16_t)((col + r0) / 2 - best);
        set_listed(disp + (size_t)row * dstep + col, v);
    }
}

// cell labels for training: mean of the valid disparities of the 4 x 4
// block when at least min_valid are valid, else NaN
__global__ __launch_bounds__(256) void k_window_label(const float* __restrict__ disp_f,
                                                      const int16_t* __restrict__ disp_i,
                                                      int64_t dstep, int Hq, int Wq, int min_valid,
                                                      float* __restrict__ label) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    float s = 0.0f;
    int cnt = 0;
    for (int dy = 0; dy < CNN_DS; ++dy)
        for (int dx = 0; dx < CNN_DS; ++dx) {
            const size_t i = (size_t)(Y * CNN_DS + dy) * dstep + X * CNN_DS + dx;
            float v;
            if (disp_f) {
                v = disp_f[i];
                if (__builtin_isnan(v)) continue;
            } else {
                if (disp_i[i] == INVALID16) continue;
                v = (float)disp_i[i];
            }
            s += v;
            ++cnt;
        }
    label[(size_t)Y * Wq + X] = cnt >= min_valid && cnt > 0 ? s / (float)cnt : __builtin_nanf("");
}

// total of the row counts (one block)
__global__ __launch_bounds__(256) void k_sum_counts(const int* __restrict__ row_count, int H,
                                                    unsigned long long* __restrict__ total) {
    __shared__ unsigned long long part[256];
    unsigned long long s = 0;
    for (int r = threadIdx.x; r < H; r += 256) s += (unsigned long long)row_count[r];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = part[0];
}

// sum of the per-cell losses and label counts, added to acc[0], acc[1]
__global__ __launch_bounds__(256) void k_loss_sum(const float* __restrict__ cell_loss, size_t cells,
                                                  float* __restrict__ acc) {
    __shared__ double sl[256], sc[256];
    double l = 0.0, c = 0.0;
    for (size_t i = threadIdx.x; i < cells; i += 256) {
        l += cell_loss[i];
        c += cell_loss[cells + i];
    }
    sl[threadIdx.x] = l;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sl[threadIdx.x] += sl[threadIdx.x + o];
            sc[threadIdx.x] += sc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] += (float)sl[0];
        acc[1] += (float)sc[0];
    }
}

}  // namespace bicos_hip
