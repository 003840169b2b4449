// Full-row Hamming argmin search, gfx950 vector-ALU form.
//
// Reference semantics (include/impl/cpu/bicos.hpp:50-113): for each query
// pixel scan all candidate columns of the same row, keep the first column
// with the strict minimum Hamming cost; with NODUPES the match is invalid if
// that minimum is attained at two or more columns.
//
// Formulation used here: every candidate gets the key (cost << 16) | column.
// Keys are unique per column, so the first strict minimum is the smallest key
// and "minimum attained twice" is "the two smallest keys share the cost
// field". The two smallest keys of a set do not depend on scan order, which
// lets the row be scanned in LDS-sized chunks. Per candidate: WORDS x
// (v_xor + v_bcnt), one v_lshl_or, one v_med3_u32 (second-smallest key, only
// with NODUPES) and one v_min_u32.
//
// Layout: a block serves one image row. The candidate row is staged in LDS
// (chunks of up to 64 KiB) and read by broadcast (every lane of a wave reads
// the same address); each lane holds R query descriptors in registers.
#pragma once

#include "device_common.hpp"

namespace bicos_hip {

enum SearchMode : int {
    SEARCH_DISP = 0,  // out = col - best (plain variants)
    SEARCH_REV = 1,   // out = best column (reverse pass of CONSISTENCY)
    SEARCH_LR = 2,    // forward pass of CONSISTENCY, checked against the reverse pass
};

constexpr int SEARCH_THREADS = 256;

template<int WORDS, bool NODUPES, int MODE, int R>
__global__ __launch_bounds__(SEARCH_THREADS) void k_search(
    const uint32_t* __restrict__ dq, const uint32_t* __restrict__ dc, int W, int chunk_cols,
    const int16_t* __restrict__ rev, int max_lr_diff, int16_t* __restrict__ out, int64_t ostep) {
    extern __shared__ uint4 lds4[];
    uint32_t* lds = (uint32_t*)lds4;

    const int row = blockIdx.y;
    const int tid = threadIdx.x;
    const int base = blockIdx.x * (SEARCH_THREADS * R);

    uint32_t q[R][WORDS];
    uint32_t b1[R], b2[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int col = base + r * SEARCH_THREADS + tid;
        const uint32_t* src = dq + ((size_t)row * W + (col < W ? col : 0)) * WORDS;
#pragma unroll
        for (int k = 0; k < WORDS; ++k) q[r][k] = src[k];
        b1[r] = 0xFFFFFFFFu;
        b2[r] = 0xFFFFFFFFu;
    }

    const uint32_t* crow = dc + (size_t)row * W * WORDS;
    for (int c0 = 0; c0 < W; c0 += chunk_cols) {
        const int cn = min(chunk_cols, W - c0);
        __syncthreads();
        if constexpr (WORDS >= 4) {
            const uint4* src = (const uint4*)(crow + (size_t)c0 * WORDS);
            for (int i = tid; i < cn * (WORDS / 4); i += SEARCH_THREADS) lds4[i] = src[i];
        } else {
            const uint32_t* src = crow + (size_t)c0 * WORDS;
            for (int i = tid; i < cn * WORDS; i += SEARCH_THREADS) lds[i] = src[i];
        }
        __syncthreads();

#pragma unroll 2
        for (int c = 0; c < cn; ++c) {
            uint32_t d[WORDS];
#pragma unroll
            for (int k = 0; k < WORDS; ++k) d[k] = lds[c * WORDS + k];
            const uint32_t ckey = (uint32_t)(c0 + c);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                uint32_t cost = 0;
#pragma unroll
                for (int k = 0; k < WORDS; ++k) cost = __popc(q[r][k] ^ d[k]) + cost;
                const uint32_t key = (cost << 16) | ckey;
                if constexpr (NODUPES) b2[r] = umed3(b1[r], b2[r], key);
                b1[r] = min(b1[r], key);
            }
        }
    }

#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int col = base + r * SEARCH_THREADS + tid;
        if (col >= W)
            continue;
        const int best = (int)(b1[r] & 0xFFFFu);
        const bool valid = !(NODUPES && (b1[r] >> 16) == (b2[r] >> 16));
        int16_t v = INVALID16;
        if (valid) {
            if constexpr (MODE == SEARCH_DISP) {
                v = (int16_t)(col - best);
            } else if constexpr (MODE == SEARCH_REV) {
                v = (int16_t)best;
            } else {
                // bicos.hpp:99-106: reverse search of the matched candidate
                const int r0 = rev[(size_t)row * W + best];
                if (r0 != INVALID16 && abs(col - r0) <= max_lr_diff)
                    v = (int16_t)((col + r0) / 2 - best);
            }
        }
        out[(size_t)row * ostep + col] = v;
    }
}

}  // namespace bicos_hip
