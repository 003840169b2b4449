// Full-row Hamming argmin search on the gfx950 matrix cores.
//
// Same result as k_search (include/impl/cpu/bicos.hpp:50-113 semantics), but
// the Hamming costs of a 32 x 32 block of (candidate, query) pairs come from
// v_mfma_i32_32x32x32_i8 instead of xor/popcount:
//
//   query q (the pixel searched for)  -> int8 entries 1 - 2 q_k  (+1 / -1)
//   candidate c                       -> int8 entries c_k        ( 0 / 1 )
//   sum_k c_k (1 - 2 q_k) = popcount(c) - 2 popcount(c & q)
//                         = popcount(c ^ q) - popcount(q)
//
// popcount(q) is constant per query, so the matrix product ranks candidates
// exactly like the Hamming cost. For rows up to 4096 columns the product is
// scaled and offset so that it is the candidate's sort key itself (cost,
// then column). The keys of a tile are reduced in pairs: smallest key with
// v_min3_i32, second smallest (no-duplicates only) with v_med3_i32 per pair
// and v_min3_i32 over the pairs, 1.25 VALU ops per output (0.5 without the
// duplicate test). Keys are unique per column, so the first strict minimum
// and the duplicate test come out as in the reference scan, independent of
// evaluation order. The int8 kernel below serves 32-bit descriptors and rows
// wider than 4096 columns; k_search_mfma_fp4 (further down) the rest.
//
// Layout. A block serves WAVES*NB*32 query pixels of one image row; each
// wave holds NB query tiles of 32 pixels as MFMA B operands in registers
// (expanded once). The candidate row is expanded to one byte per descriptor
// bit and staged in LDS in chunks of CH columns, padded by 16 bytes per
// column so the 128-bit operand reads are bank-conflict free. Output lane
// layout of the 32x32 MFMA: lane l, accumulator i holds query column
// n = l % 32 and candidate row m = 8 (i / 4) + 4 (l / 32) + i % 4. Lanes l and
// l ^ 32 see the same query and disjoint candidates; their minima are merged
// at the end.
#pragma once

#include "device_common.hpp"
#include "search.hpp"

namespace bicos_hip {


typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// v_mul_u32_u24 (full rate; the 32-bit multiply is quarter rate). Written
// out because the compiler folds a following shift into the constant, which
// then no longer fits 24 bits and takes the 32-bit multiply.
__device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// 4 bits -> 4 bytes of 0/1 (bit b to byte b): the nibble times 0x00204081
// puts bit b at 8 b
__device__ __forceinline__ uint32_t spread4(uint32_t x) {
    return mul_u24(x & 0xFu, 0x00204081u) & 0x01010101u;
}

// 16 descriptor bits -> 16 int8 entries of 0/1
__device__ __forceinline__ uint4 expand01(uint32_t b16) {
    return make_uint4(spread4(b16), spread4(b16 >> 4), spread4(b16 >> 8), spread4(b16 >> 12));
}

// 16 descriptor bits -> 16 int8 entries of +1 (bit 0) / -1 (bit 1)
__device__ __forceinline__ v4i expand_pm1(uint32_t b16) {
    v4i r;
    r[0] = (int)(spread4(b16) * 0xFEu | 0x01010101u);
    r[1] = (int)(spread4(b16 >> 4) * 0xFEu | 0x01010101u);
    r[2] = (int)(spread4(b16 >> 8) * 0xFEu | 0x01010101u);
    r[3] = (int)(spread4(b16 >> 12) * 0xFEu | 0x01010101u);
    return r;
}

// 16 descriptor bits -> 16 int8 entries of 0 / 64 (embedded-key operands)
__device__ __forceinline__ uint4 expand0_64(uint32_t b16) {
    const uint4 v = expand01(b16);
    return make_uint4(v.x << 6, v.y << 6, v.z << 6, v.w << 6);
}

// 16 descriptor bits -> 16 int8 entries of +64 (bit 0) / -64 (bit 1)
__device__ __forceinline__ v4i expand_pm64(uint32_t b16) {
    v4i r;
    r[0] = (int)((spread4(b16) << 7) | 0x40404040u);
    r[1] = (int)((spread4(b16 >> 4) << 7) | 0x40404040u);
    r[2] = (int)((spread4(b16 >> 8) << 7) | 0x40404040u);
    r[3] = (int)((spread4(b16 >> 12) << 7) | 0x40404040u);
    return r;
}

__device__ __forceinline__ int imin3(int a, int b, int c) {
    // lowered to v_min3_i32
    return min(min(a, b), c);
}

// The pair step of the no-duplicates reduction, t = med3(s, a, b) and
// s' = min3(s, a, b), as two instructions: in C the two share a two-operand
// min after the compiler's rewriting and take three. Inline asm gets no
// MFMA-result hazard wait states from the compiler, so a and b must not be
// the first reads of a matrix product's result: callers order this after a
// compiler-visible read of the same product (through s).
__device__ __forceinline__ void pair_step_asm(int& s, int a, int b, int& t) {
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(t) : "v"(s), "v"(a), "v"(b));
    asm("v_min3_i32 %0, %1, %2, %3" : "=v"(s) : "v"(s), "v"(a), "v"(b));
}

// The first pair of a tile: s' = min3(s, a, b) in C (the product's first
// read, which gets the hazard wait states) and t = med3(s, a, b) written out
// with s' as an extra, unused input, so that it cannot be scheduled before
// that read.
__device__ __forceinline__ void pair_step_first(int& s, int a, int b, int& t) {
    const int s_new = imin3(s, a, b);
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(t) : "v"(s), "v"(a), "v"(b), "v"(s_new));
    s = s_new;
}

// "no candidate" key: above every real key (|key| < 2^21 in both key forms)
// and far enough below 2^31 that the state offsets (at most 32 * 1024 tiles)
// cannot wrap it around.
constexpr int NO_KEY = 0x3FFFFFFF;

// Largest row width of the embedded-key form: the column field is 12 bits.
constexpr int EMBED_MAX_W = 4096;

// Keys. A key orders candidates by (cost, column) and is unique per column.
//
// EMBED = false: key = ((dot + 256) << 16) + column, one v_lshl_add_u32 per
//   output after the products (query entries +1 / -1, candidate entries 0 / 1).
// EMBED = true (rows up to 4096 columns): the matrix product itself yields
//   the key. Query entries are +64 / -64 and candidate entries 0 / 64, so the
//   product is 4096 * dot; the accumulator of the first K step starts from the
//   candidate's row within the 32-column tile instead of 0. The sum is
//   4096 * dot + row, exact in int32, with no VALU work per output before
//   the reduction.
//
// In both forms the running state is kept relative to the current tile
// (minus 32 per tile), so the tile base never has to be added per output.
//
// LIST = true: the queries of row r are the listed columns
// qcol[r * W + i], i < row_count[r] (negative entries are skipped); the row
// is blockIdx.x so that rows, most of which need one block, spread over the
// XCDs. Results go to out[r * ostep + col] (SEARCH_DISP) or to
// res[r * W + i] (SEARCH_REV). Waves without listed queries skip the
// arithmetic but take part in the staging.
template<int WORDS, int NB, bool NODUPES, int MODE, int WAVES, int CH, bool LIST = false,
         bool EMBED = false>
__global__ __launch_bounds__(WAVES * 64) void k_search_mfma(
    const uint32_t* __restrict__ dq, const uint32_t* __restrict__ dc, int W,
    const int16_t* __restrict__ rev, int max_lr_diff, int16_t* __restrict__ out, int64_t ostep,
    const int16_t* __restrict__ qcol, const int* __restrict__ row_count, int16_t* __restrict__ res) {
    constexpr int KS = WORDS;              // K steps of 32 bits
    constexpr int STRIDE = 2 * KS + 1;     // uint4 per staged candidate (16 B pad)
    static_assert(CH % 32 == 0, "chunks of whole 32-column tiles (staging rounds a chunk up to a tile)");
    constexpr int CSHIFT = EMBED ? 12 : 16;           // column field width of a key
    constexpr int CMASK = (1 << CSHIFT) - 1;
    extern __shared__ uint4 lds[];

    int row, group;
    if constexpr (LIST) {
        row = blockIdx.x;
        group = blockIdx.y;
    } else {
        xcd_row_group(row, group);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nl = lane & 31, g = lane >> 5;
    int count = 0;
    if constexpr (LIST) {
        count = row_count[row];
        if (group * WAVES * NB * 32 >= count)
            return;  // uniform over the block
    }

    // ---- query operands: bits 32 s + 16 g .. +15 of each query pixel
    v4i qop[NB][KS];
    int s1[NB], s2[NB];
    int qc[NB];
    const int qbase = (group * WAVES + wave) * (NB * 32);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int col = qbase + b * 32 + nl;
        if constexpr (LIST) col = col < count ? qcol[(size_t)row * W + col] : -1;
        qc[b] = col;
        const uint32_t* src = dq + ((size_t)row * W + (col >= 0 && col < W ? col : 0)) * WORDS;
#pragma unroll
        for (int s = 0; s < KS; ++s)
            qop[b][s] = EMBED ? expand_pm64(src[s] >> (16 * g)) : expand_pm1(src[s] >> (16 * g));
        s1[b] = NO_KEY;
        s2[b] = NO_KEY;
    }
    // waves whose query tiles lie past the row (or the list) skip the
    // arithmetic but take part in the staging
    const bool busy = qbase < (LIST ? count : W);

    // candidate row of accumulator i within a tile (the C operand of the
    // first K step in the embedded form)
    v16i crow_init;
#pragma unroll
    for (int i = 0; i < 16; ++i) crow_init[i] = EMBED ? 8 * (i / 4) + 4 * g + (i % 4) : 0;

    // Staging: a thread expands whole candidate columns (CPT per chunk). The
    // raw words of the next chunk are loaded before the current chunk's
    // tiles, so their latency hides behind the matrix-core work; columns past
    // the row stage as zero bits.
    constexpr int NT = WAVES * 64;
    constexpr int CPT = (CH + NT - 1) / NT;
    const uint32_t* crow = dc + (size_t)row * W * WORDS;
    uint32_t raw[CPT][WORDS];
    auto fetch = [&](int c0) {
        const int cn = min(CH, W - c0);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int m = (int)threadIdx.x + j * NT;
#pragma unroll
            for (int w = 0; w < WORDS; ++w)
                raw[j][w] = m < cn ? crow[(size_t)(c0 + m) * WORDS + w] : 0u;
        }
    };
    fetch(0);
    const int ntiles = (W + 31) / 32;
    for (int c0 = 0; c0 < W; c0 += CH) {
        const int cn = min(CH, W - c0);
        const int cnp = (cn + 31) & ~31;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int m = (int)threadIdx.x + j * NT;
            if (m < cnp) {
#pragma unroll
                for (int c = 0; c < 2 * KS; ++c) {
                    const uint32_t bits = (raw[j][c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
                    lds[m * STRIDE + c] = EMBED ? expand0_64(bits) : expand01(bits);
                }
            }
        }
        __syncthreads();
        if (c0 + CH < W) fetch(c0 + CH);

        // full tiles; the NB product chains are independent, so matrix-core
        // work of one overlaps the VALU work of the previous
        const int nfull = busy ? cn / 32 : 0;
        for (int mt = 0; mt < nfull; ++mt) {
            v4i aop[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint4 v = lds[(mt * 32 + nl) * STRIDE + 2 * s + g];
                aop[s] = v4i{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
            }
            v16i acc[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[0], qop[b][0], crow_init, 0, 0, 0);
#pragma unroll
                for (int s = 1; s < KS; ++s)
                    acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[s], qop[b][s], acc[b], 0, 0, 0);
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                // keys in pairs (x0, x1): with s1 <= s2 the two smallest so
                // far, the new smallest is min3(s1, x0, x1) and the new
                // second smallest min(s2, med3(s1, x0, x1)); the eight med3
                // values of a tile fold into s2 with four min3. 1.25 VALU
                // ops per output (0.5 without the duplicate test).
                int t[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    // candidate row within the tile of output i:
                    // 8 (i / 4) + 4 g + i % 4; the tile base (and 4 g in the
                    // shifted form) live in the running state offset
                    int x[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int i = 2 * p + h;
                        x[h] = EMBED ? acc[b][i]
                                     : (int)(((uint32_t)acc[b][i] << 16) +
                                             ((256u << 16) + 8u * (i / 4) + (i % 4)));
                    }
                    if constexpr (NODUPES) {
                        // every pair after the first depends on the first
                        // through s1
                        if (p == 0)
                            pair_step_first(s1[b], x[0], x[1], t[p]);
                        else
                            pair_step_asm(s1[b], x[0], x[1], t[p]);
                    } else {
                        s1[b] = imin3(s1[b], x[0], x[1]);
                    }
                }
                if constexpr (NODUPES)
                    s2[b] = imin3(imin3(s2[b], t[0], t[1]), imin3(t[2], t[3], t[4]), imin3(t[5], t[6], t[7])) - 32;
                s1[b] -= 32;
            }
        }
        // ragged last tile of the row: candidates at or beyond W are masked
        if (busy && nfull * 32 < cn) {
            const int mt = nfull;
            v4i aop[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint4 v = lds[(mt * 32 + nl) * STRIDE + 2 * s + g];
                aop[s] = v4i{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
            }
            const int tbase = c0 + mt * 32;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                v16i acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[0], qop[b][0], crow_init, 0, 0, 0);
#pragma unroll
                for (int s = 1; s < KS; ++s)
                    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[s], qop[b][s], acc, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    int key = EMBED ? acc[i]
                                    : (int)(((uint32_t)acc[i] << 16) +
                                            ((256u << 16) + 8u * (i / 4) + (i % 4)));
                    if (tbase + 8 * (i / 4) + 4 * g + (i % 4) >= W) key = NO_KEY;
                    if constexpr (NODUPES) s2[b] = imed3(s1[b], s2[b], key);
                    s1[b] = min(s1[b], key);
                }
                s1[b] -= 32;
                if constexpr (NODUPES) s2[b] -= 32;
            }
        }
    }

    // restore absolute keys: + 32 * (tiles scanned) (+ 4 g in the shifted
    // form), then merge lanes l, l ^ 32
    const int off = 32 * ntiles + (EMBED ? 0 : 4 * g);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int b1 = s1[b] + off, b2 = s2[b] + off;
        const int o1 = __shfl_xor(b1, 32);
        const int o2 = __shfl_xor(b2, 32);
        const int m1 = min(b1, o1);
        const int m2 = min(max(b1, o1), min(b2, o2));
        const int col = qc[b];
        if constexpr (LIST) {
            const int i = qbase + b * 32 + nl;
            if (g != 0 || i >= count)
                continue;
            const bool valid = col >= 0 && !(NODUPES && (m1 >> CSHIFT) == (m2 >> CSHIFT));
            const int16_t best = valid ? (int16_t)(m1 & CMASK) : INVALID16;
            if constexpr (MODE == SEARCH_REV)
                res[(size_t)row * W + i] = best;
            else if (col >= 0)
                set_listed(out + (size_t)row * ostep + col, valid ? (int16_t)(col - best) : INVALID16);
            continue;
        }
        if (g != 0 || col >= W)
            continue;
        const int best = m1 & CMASK;
        const bool valid = !(NODUPES && (m1 >> CSHIFT) == (m2 >> CSHIFT));
        int16_t v = INVALID16;
        if (valid) {
            if constexpr (MODE == SEARCH_DISP) {
                v = (int16_t)(col - best);
            } else if constexpr (MODE == SEARCH_REV) {
                v = (int16_t)best;
            } else {
                const int r0 = rev[(size_t)row * W + best];
                if (r0 != INVALID16 && abs(col - r0) <= max_lr_diff)
                    v = (int16_t)((col + r0) / 2 - best);
            }
        }
        out[(size_t)row * ostep + col] = v;
    }
}

// ---------------------------------------------------------------------------
// fp4 form (rows up to 4096 columns, descriptors of 64 bits and more).
//
// v_mfma_scale_f32_32x32x64_f8f6f4 with both operands in fp4 (E2M1) runs at
// twice the int8 rate and covers 64 descriptor bits per product. Candidate
// bits become 1.0 / 0 (codes 0b0010 / 0), query bits +1.0 / -1.0 (codes
// 0b0010 / 0b1010); the E8M0 block scales 2^7 (candidates) and 2^6
// (queries) make the product 8192 * dot, with dot = popcount(c) - 2
// popcount(c & q) as above. All values are integers below 2^24, so the FP32
// accumulation is exact. Lane l of an
// operand holds row (column) l % 32; descriptor word 2 s + g of K step s
// goes to lane half g, its bit pair p (bits 2 p and 2 p + 1), p = 4 i + r,
// to byte i of register r, the lower bit in the low nibble (the same packing
// for both operands, which is all the product needs).
//
// Keys are FP32 values 2^22 + 8192 * dot + column (the accumulator starts
// from 2^22 + the row within the tile; |8192 dot| <= 2^21), all positive, so
// their bit patterns order like the values and the integer v_med3_i32 /
// v_min3_i32 reduction is unchanged. The running minimum is kept relative to
// the current tile (an exact FP32 subtraction of 32 per tile), so that its
// column comes out exact. The second smallest key is only asked for its cost
// (the duplicate test) and is left in the frame of the tile it came from: its
// column then reads up to 32 * 128 = 4096 too large, which stays below one
// cost step of 8192, so no comparison between keys of different costs
// changes and its cost comes out exact.
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int KEY4_BASE = 1 << 22;
constexpr int KEY4_NONE = 0x4B000000;   // bit pattern of 2^23: above every real key

// 32 descriptor bits -> 32 fp4 entries in the packing above: v_perm_b32
// looks every bit pair up in a table of 4 bytes (the selector byte is the
// pair's value 0..3, which picks byte 0..3 of the table), 11 VALU
// instructions per word.
template<uint32_t TABLE>
__device__ __forceinline__ uint4 expand_fp4(uint32_t w) {
    constexpr uint32_t M = 0x03030303u;
    return make_uint4(__builtin_amdgcn_perm(TABLE, TABLE, w & M),
                      __builtin_amdgcn_perm(TABLE, TABLE, (w >> 2) & M),
                      __builtin_amdgcn_perm(TABLE, TABLE, (w >> 4) & M),
                      __builtin_amdgcn_perm(TABLE, TABLE, (w >> 6) & M));
}

// candidate bits -> 1.0 / 0: pair values 0, 1, 2, 3 -> bytes 0x00, 0x02,
// 0x20, 0x22
__device__ __forceinline__ uint4 expand_c_fp4(uint32_t w) { return expand_fp4<0x22200200u>(w); }

// query bits -> +1.0 (bit 0) / -1.0 (bit 1): pair values 0, 1, 2, 3 ->
// bytes 0x22, 0x2A, 0xA2, 0xAA
__device__ __forceinline__ v8i expand_q_fp4(uint32_t w) {
    const uint4 v = expand_fp4<0xAAA22A22u>(w);
    return v8i{(int)v.x, (int)v.y, (int)v.z, (int)v.w, 0, 0, 0, 0};
}

__device__ __forceinline__ int f_sub32(int key) {
    return __float_as_int(__int_as_float(key) - 32.0f);
}

template<int WORDS, int NB, bool NODUPES, int MODE, int WAVES, int CH, bool LIST = false>
__global__ __launch_bounds__(WAVES * 64) void k_search_mfma_fp4(
    const uint32_t* __restrict__ dq, const uint32_t* __restrict__ dc, int W,
    const int16_t* __restrict__ rev, int max_lr_diff, int16_t* __restrict__ out, int64_t ostep,
    const int16_t* __restrict__ qcol, const int* __restrict__ row_count, int16_t* __restrict__ res) {
    static_assert(WORDS % 2 == 0, "fp4 form: 64-bit K steps");
    static_assert(CH % 32 == 0, "chunks of whole 32-column tiles (staging rounds a chunk up to a tile)");
    constexpr int KS = WORDS / 2;          // K steps of 64 bits
    constexpr int STRIDE = WORDS + 1;      // uint4 per staged candidate (16 B pad)
    constexpr int SCALE_C = 127 + 7;       // E8M0 2^7 for the candidates,
    constexpr int SCALE_Q = 127 + 6;       // 2^6 for the queries
    extern __shared__ uint4 lds[];

    int row, group0;
    if constexpr (LIST) {
        row = blockIdx.x;
        group0 = blockIdx.y;
    } else {
        xcd_row_group(row, group0);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nl = lane & 31, g = lane >> 5;
    // LIST: a block serves groups blockIdx.y, blockIdx.y + gridDim.y, ...
    // of WAVES * NB * 32 entries of its row (staging the row again per group)
    const int count = LIST ? row_count[row] : 0;
    for (int group = group0;; group += gridDim.y) {
    if constexpr (LIST) {
        if (group * WAVES * NB * 32 >= count)
            return;  // uniform over the block
    }

    v8i qop[NB][KS];
    int s1[NB], s2[NB];
    int qc[NB];
    const int qbase = (group * WAVES + wave) * (NB * 32);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int col = qbase + b * 32 + nl;
        if constexpr (LIST) col = col < count ? qcol[(size_t)row * W + col] : -1;
        qc[b] = col;
        const uint32_t* src = dq + ((size_t)row * W + (col >= 0 && col < W ? col : 0)) * WORDS;
#pragma unroll
        for (int s = 0; s < KS; ++s) qop[b][s] = expand_q_fp4(src[2 * s + g]);
        s1[b] = KEY4_NONE;
        s2[b] = KEY4_NONE;
    }
    const bool busy = qbase < (LIST ? count : W);

    v16f crow_init;
#pragma unroll
    for (int i = 0; i < 16; ++i) crow_init[i] = (float)(KEY4_BASE + 8 * (i / 4) + 4 * g + (i % 4));

    constexpr int NT = WAVES * 64;
    constexpr int CPT = (CH + NT - 1) / NT;
    const uint32_t* crow = dc + (size_t)row * W * WORDS;
    uint32_t raw[CPT][WORDS];
    auto fetch = [&](int c0) {
        const int cn = min(CH, W - c0);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int m = (int)threadIdx.x + j * NT;
#pragma unroll
            for (int w = 0; w < WORDS; ++w)
                raw[j][w] = m < cn ? crow[(size_t)(c0 + m) * WORDS + w] : 0u;
        }
    };
    fetch(0);
    const int ntiles = (W + 31) / 32;
    for (int c0 = 0; c0 < W; c0 += CH) {
        const int cn = min(CH, W - c0);
        const int cnp = (cn + 31) & ~31;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int m = (int)threadIdx.x + j * NT;
            if (m < cnp) {
#pragma unroll
                for (int w = 0; w < WORDS; ++w) lds[m * STRIDE + w] = expand_c_fp4(raw[j][w]);
            }
        }
        __syncthreads();
        if (c0 + CH < W) fetch(c0 + CH);

        // full tiles, then the ragged last tile of the row (candidates at or
        // beyond W masked)
        const int nfull = busy ? cn / 32 : 0;
        auto products = [&](int mt, v16f* acc) {
            v8i aop[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint4 v = lds[(mt * 32 + nl) * STRIDE + 2 * s + g];
                aop[s] = v8i{(int)v.x, (int)v.y, (int)v.z, (int)v.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                acc[b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aop[0], qop[b][0], crow_init, 4, 4,
                                                                         0, SCALE_C, 0, SCALE_Q);
#pragma unroll
                for (int s = 1; s < KS; ++s)
                    acc[b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aop[s], qop[b][s], acc[b], 4, 4,
                                                                             0, SCALE_C, 0, SCALE_Q);
            }
        };
        for (int mt = 0; mt < nfull; ++mt) {
            v16f acc[NB];
            products(mt, acc);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                int t[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    const int x0 = __float_as_int(acc[b][2 * p]), x1 = __float_as_int(acc[b][2 * p + 1]);
                    if constexpr (NODUPES) {
                        // every pair after the first depends on the first
                        // through s1
                        if (p == 0)
                            pair_step_first(s1[b], x0, x1, t[p]);
                        else
                            pair_step_asm(s1[b], x0, x1, t[p]);
                    } else {
                        s1[b] = imin3(s1[b], x0, x1);
                    }
                }
                if constexpr (NODUPES)
                    s2[b] = imin3(imin3(s2[b], t[0], t[1]), imin3(t[2], t[3], t[4]),
                                  imin3(t[5], t[6], t[7]));
                s1[b] = f_sub32(s1[b]);
            }
        }
        if (busy && nfull * 32 < cn) {
            const int tbase = c0 + nfull * 32;
            v16f acc[NB];
            products(nfull, acc);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    int key = __float_as_int(acc[b][i]);
                    if (tbase + 8 * (i / 4) + 4 * g + (i % 4) >= W) key = KEY4_NONE;
                    if constexpr (NODUPES) s2[b] = imed3(s1[b], s2[b], key);
                    s1[b] = min(s1[b], key);
                }
                s1[b] = f_sub32(s1[b]);
            }
        }
    }

    // absolute keys 8192 * dot + column (the second one's column read up to
    // 4096 too large), then merge lanes l, l ^ 32
    const float off = (float)(32 * ntiles - KEY4_BASE);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int b1 = (int)(__int_as_float(s1[b]) + off), b2 = (int)(__int_as_float(s2[b]) + off);
        const int o1 = __shfl_xor(b1, 32);
        const int o2 = __shfl_xor(b2, 32);
        const int m1 = min(b1, o1);
        const int m2 = min(max(b1, o1), min(b2, o2));
        const int col = qc[b];
        const bool valid_key = !(NODUPES && (m1 >> 13) == (m2 >> 13));
        const int best = m1 & 8191;
        if constexpr (LIST) {
            const int i = qbase + b * 32 + nl;
            if (g != 0 || i >= count)
                continue;
            const bool valid = col >= 0 && valid_key;
            const int16_t bst = valid ? (int16_t)best : INVALID16;
            if constexpr (MODE == SEARCH_REV)
                res[(size_t)row * W + i] = bst;
            else if (col >= 0)
                set_listed(out + (size_t)row * ostep + col, valid ? (int16_t)(col - best) : INVALID16);
            continue;
        }
        if (g != 0 || col >= W)
            continue;
        int16_t v = INVALID16;
        if (valid_key) {
            if constexpr (MODE == SEARCH_DISP) {
                v = (int16_t)(col - best);
            } else if constexpr (MODE == SEARCH_REV) {
                v = (int16_t)best;
            } else {
                const int r0 = rev[(size_t)row * W + best];
                if (r0 != INVALID16 && abs(col - r0) <= max_lr_diff)
                    v = (int16_t)((col + r0) / 2 - best);
            }
        }
        out[(size_t)row * ostep + col] = v;
    }
    if constexpr (!LIST) return;
    __syncthreads();  // the next group stages the row again
    }
}

}  // namespace bicos_hip
