// Temporal binary descriptor transform, gfx950.
//
// One lane per pixel; lanes of a wave cover consecutive columns, so every
// frame-plane read is a contiguous 64- or 128-byte span (coalesced). The n
// samples of a pixel stay in registers; the bit emission order is the
// reference's (include/impl/cpu/descriptor_transform.hpp:31-123) with each
// bit's position resolved at compile time.
#pragma once

#include "device_common.hpp"

namespace bicos_hip {

template<int WORDS>
__device__ __forceinline__ void setbit(uint32_t (&w)[WORDS], int pos, bool v) {
    w[pos >> 5] |= (uint32_t)v << (pos & 31);
}

template<int WORDS>
__device__ __forceinline__ void store_desc(uint32_t* __restrict__ out, const uint32_t (&w)[WORDS],
                                           int out_words) {
    if (out_words == WORDS) {
        if constexpr (WORDS >= 4) {
#pragma unroll
            for (int k = 0; k < WORDS; k += 4)
                *(uint4*)(out + k) = make_uint4(w[k], w[k + 1], w[k + 2], w[k + 3]);
        } else if constexpr (WORDS == 2) {
            *(uint2*)out = make_uint2(w[0], w[1]);
        } else {
            out[0] = w[0];
        }
    } else {
        for (int k = 0; k < out_words; ++k)
            out[k] = k < WORDS ? w[k] : 0u;
    }
}

// Limited mode (descriptor_transform.hpp:31-73, descriptor_transform.cuh:91-149).
// n is a runtime value <= NMAX. Bits of frame triple t sit at fixed positions
// (3t for t < 2, 4t - 2 after that, when the pair-sum bit joins); the four
// closing bits go to 3(n-2) for n < 4 and to 4n - 10 otherwise.
template<typename TIn, int WORDS, int NMAX>
__global__ __launch_bounds__(256) void k_transform_limited(const Stack f, int n, int W,
                                                           int out_words,
                                                           uint32_t* __restrict__ out) {
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;

    uint32_t p[NMAX];
    float av = 0.0f;
#pragma unroll
    for (int t = 0; t < NMAX; ++t) {
        p[t] = 0;
        if (t < n) {
            p[t] = load_sample_cached<TIn>(f, t, row, col);
            av += (float)p[t];
        }
    }
    av /= (float)n;

    uint32_t w[WORDS];
#pragma unroll
    for (int k = 0; k < WORDS; ++k) w[k] = 0u;

    int prev0 = -1, prev1 = -1;
#pragma unroll
    for (int t = 0; t < NMAX - 2; ++t) {
        if (t < n - 2) {
            const uint32_t a = p[t], b = p[t + 1], c = p[t + 2];
            const int pos = t < 2 ? 3 * t : 4 * t - 2;
            setbit(w, pos + 0, a < b);
            setbit(w, pos + 1, a < c);
            setbit(w, pos + 2, (float)a < av);
            const int cur = (int)(a + b);
            if (t >= 2)
                setbit(w, pos + 3, ((t & 1) ? prev1 : prev0) < cur);
            if (t & 1) prev1 = cur; else prev0 = cur;
        }
    }

    uint32_t a = 0, b = 0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t) {
        if (t == n - 2) a = p[t];
        if (t == n - 1) b = p[t];
    }
    const int prevf = ((n - 2) & 1) ? prev1 : prev0;
    const uint32_t fb = (uint32_t)(a < b) | ((uint32_t)((float)a < av) << 1)
                      | ((uint32_t)((float)b < av) << 2) | ((uint32_t)(prevf < (int)(a + b)) << 3);
    const int P = n >= 4 ? 4 * n - 10 : 3 * (n - 2);
    const int wi = P >> 5, sh = P & 31;
#pragma unroll
    for (int k = 0; k < WORDS; ++k) {
        if (k == wi) w[k] |= fb << sh;
        if (k == wi + 1 && sh > 28) w[k] |= fb >> (32 - sh);
    }

    store_desc(out + ((size_t)row * W + col) * out_words, w, out_words);
}

// Full mode (descriptor_transform.hpp:75-123, descriptor_transform.cuh:30-89)
// for a compile-time stack depth N (2..16).
template<typename TIn, int N>
__global__ __launch_bounds__(256) void k_transform_full(const Stack f, int W, int out_words,
                                                        uint32_t* __restrict__ out) {
    constexpr int BITS = N * N - 2 * N + 3;
    constexpr int WORDS = BITS <= 32 ? 1 : BITS <= 64 ? 2 : BITS <= 128 ? 4 : 8;
    const int col = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    if (col >= W)
        return;

    uint32_t p[N];
    float av = 0.0f;
#pragma unroll
    for (int t = 0; t < N; ++t) {
        p[t] = load_sample_cached<TIn>(f, t, row, col);
        av += (float)p[t];
    }
    av /= (float)N;

    uint32_t w[WORDS];
#pragma unroll
    for (int k = 0; k < WORDS; ++k) w[k] = 0u;

    uint32_t s[N - 1];  // pair sums in the wider type (impl/common.hpp:28-42)
    int pos = 0;
#pragma unroll
    for (int t = 0; t < N - 2; ++t) {
        const uint32_t a = p[t], b = p[t + 1], c = p[t + 2];
        setbit(w, pos++, a < b);
        setbit(w, pos++, a < c);
        setbit(w, pos++, (float)a < av);
        s[t] = a + b;
    }
    s[N - 2] = p[N - 2] + p[N - 1];
    setbit(w, pos++, p[N - 2] < p[N - 1]);
    setbit(w, pos++, (float)p[N - 2] < av);
    setbit(w, pos++, (float)p[N - 1] < av);
#pragma unroll
    for (int t = 0; t < N - 1; ++t)
#pragma unroll
        for (int i = 0; i < N - 1; ++i) {
            if (i == t || i == t - 1 || i == t + 1)
                continue;
            setbit(w, pos++, s[t] < s[i]);
        }

    store_desc(out + ((size_t)row * W + col) * out_words, w, out_words);
}

// ---------------------------------------------------------------------------
// Vectorised form: each lane transforms PX = 4 adjacent pixels, reading 4
// samples per frame with one 8-byte (u16) or 4-byte (u8) load. Samples are
// kept as floats (exact: integers below 2^24), so every comparison
// `x < y` is the sign bit of x - y. Bits of one emission group are shifted in
// with v_alignbit (w = (w << 1) | (d >> 31)), last emission first, so the
// group lands in natural order and is OR-ed into its word at a compile-time
// position. blockIdx.z selects the view (0: out0 from f0, 1: out1 from f1).
// Requires 8-byte (u16) / 4-byte (u8) aligned frame rows; the host checks.

constexpr int TPX = 4;

template<typename TIn>
__device__ __forceinline__ void load_px4(const Stack& f, int t, int row, int col, int W,
                                         float (&v)[TPX]) {
    const TIn* p = frame_row<TIn>(f, t, row) + col;
    if (col + TPX <= W) {
        if constexpr (sizeof(TIn) == 2) {
            const uint2 u = *(const uint2*)p;
            v[0] = (float)(u.x & 0xFFFFu); v[1] = (float)(u.x >> 16);
            v[2] = (float)(u.y & 0xFFFFu); v[3] = (float)(u.y >> 16);
        } else {
            const uint32_t u = *(const uint32_t*)p;
            v[0] = (float)(u & 0xFFu); v[1] = (float)((u >> 8) & 0xFFu);
            v[2] = (float)((u >> 16) & 0xFFu); v[3] = (float)(u >> 24);
        }
    } else {
#pragma unroll
        for (int k = 0; k < TPX; ++k) v[k] = col + k < W ? (float)p[k] : 0.0f;
    }
}

// PX adjacent raw samples of frame t packed into 32-bit words (zero beyond
// the row end); PX * sizeof(TIn) is 4 or 8 bytes
template<typename TIn, int PX>
__device__ __forceinline__ void load_raw(const Stack& f, int t, int row, int col, int W,
                                         uint32_t (&r)[PX * sizeof(TIn) / 4]) {
    static_assert(PX * sizeof(TIn) == 4 || PX * sizeof(TIn) == 8, "4- or 8-byte groups");
    const TIn* p = frame_row<TIn>(f, t, row) + col;
    if (col + PX <= W) {
        if constexpr (PX * sizeof(TIn) == 8) {
            const uint2 u = *(const uint2*)p;
            r[0] = u.x;
            r[1] = u.y;
        } else {
            r[0] = *(const uint32_t*)p;
        }
    } else {
#pragma unroll
        for (int j = 0; j < (int)(PX * sizeof(TIn) / 4); ++j) r[j] = 0u;
#pragma unroll
        for (int k = 0; k < PX; ++k)
            if (col + k < W) {
                constexpr int PER = 4 / sizeof(TIn);
                r[k / PER] |= (uint32_t)p[k] << (8 * sizeof(TIn) * (k % PER));
            }
    }
}

template<typename TIn, int PX>
__device__ __forceinline__ void unpack(const uint32_t (&r)[PX * sizeof(TIn) / 4], float (&v)[PX]) {
    constexpr int PER = 4 / sizeof(TIn);
    constexpr uint32_t MASK = sizeof(TIn) == 2 ? 0xFFFFu : 0xFFu;
#pragma unroll
    for (int k = 0; k < PX; ++k) v[k] = (float)((r[k / PER] >> (8 * sizeof(TIn) * (k % PER))) & MASK);
}

// shift the comparison x < y into the low end of g
__device__ __forceinline__ uint32_t shin(uint32_t g, float x, float y) {
    return __builtin_amdgcn_alignbit(g, __float_as_uint(x - y), 31);
}

template<int WORDS>
__device__ __forceinline__ void put_group(uint32_t (&w)[WORDS], int pos, int len, uint32_t g) {
    // pos, len compile-time after unrolling; a group may straddle two words
    w[pos >> 5] |= g << (pos & 31);
    if ((pos & 31) + len > 32 && (pos >> 5) + 1 < WORDS)
        w[(pos >> 5) + 1] |= g >> (32 - (pos & 31));
}

template<int WORDS, int PX = TPX>
__device__ __forceinline__ void store_px(uint32_t* __restrict__ out, int W, int row, int col,
                                         const uint32_t (&w)[PX][WORDS], int out_words) {
#pragma unroll
    for (int k = 0; k < PX; ++k)
        if (col + k < W)
            store_desc(out + ((size_t)row * W + col + k) * out_words, w[k], out_words);
}

// Sample k of the PX packed raw samples of one frame
template<typename TIn, int PX>
__device__ __forceinline__ int raw_sample(const uint32_t (&r)[PX * sizeof(TIn) / 4], int k) {
    constexpr int PER = 4 / sizeof(TIn);
    constexpr uint32_t MASK = sizeof(TIn) == 2 ? 0xFFFFu : 0xFFu;
    return (int)((r[k / PER] >> (8 * sizeof(TIn) * (k % PER))) & MASK);
}

// shift the comparison x < y of integers into the low end of g
__device__ __forceinline__ uint32_t shin_i(uint32_t g, int x, int y) {
    return __builtin_amdgcn_alignbit(g, (uint32_t)(x - y), 31);
}

// limited mode (descriptor_transform.hpp:31-73); n = NE when NE > 0 (every
// frame guard resolved at compile time), else runtime n <= NMAX. PX pixels
// per lane (4, or 2 for deep 16-bit stacks, where holding 4 pixels' frames
// would cost the occupancy that hides the memory latency).
//
// All comparisons are integer: samples and pair sums are small integers,
// and for an integer p and the reference's FP32 mean av >= 0,
// p < av <=> p < ceil(av). av itself is float(sum) / n with the sum formed
// exactly (below 2^24 for n <= 65, as the reference's sequential FP32 sum
// is). Each comparison x < y is the sign bit of x - y.
template<typename TIn, int WORDS, int NMAX, int PX = TPX, int NE = 0>
__global__ __launch_bounds__(256) void k_transform_limited_v4(const Stack f0, const Stack f1,
                                                              int n_rt, int W, int out_words,
                                                              uint32_t* __restrict__ out0,
                                                              uint32_t* __restrict__ out1) {
    static_assert(NE == 0 || NE <= NMAX, "exact depth beyond NMAX");
    constexpr int RW = PX * sizeof(TIn) / 4;
    const int n = NE > 0 ? NE : n_rt;
    const int col = (blockIdx.x * 256 + threadIdx.x) * PX, row = blockIdx.y;
    if (col >= W)
        return;
    const Stack& f = blockIdx.z ? f1 : f0;
    uint32_t* __restrict__ out = blockIdx.z ? out1 : out0;

    // All frames are loaded once and held as raw samples (packed, two per
    // register for 16-bit input); the comparisons read them in place.
    uint32_t raw[NMAX][RW];
    int sum[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) sum[k] = 0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t) {
#pragma unroll
        for (int j = 0; j < RW; ++j) raw[t][j] = 0u;
        if (t < n) {
            load_raw<TIn, PX>(f, t, row, col, W, raw[t]);
#pragma unroll
            for (int k = 0; k < PX; ++k) sum[k] += raw_sample<TIn, PX>(raw[t], k);
        }
    }
    int avc[PX];  // ceil of the reference's FP32 mean
#pragma unroll
    for (int k = 0; k < PX; ++k) avc[k] = (int)__builtin_ceilf((float)sum[k] / (float)n);

    uint32_t w[PX][WORDS];
#pragma unroll
    for (int k = 0; k < PX; ++k)
#pragma unroll
        for (int j = 0; j < WORDS; ++j) w[k][j] = 0u;

    int prev0[PX], prev1[PX];  // pair sums S(t - 2) by parity of t
    int pa[PX], pb[PX], pc[PX];  // samples of frames t, t + 1, t + 2 (unpacked as they are needed)
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        pa[k] = raw_sample<TIn, PX>(raw[0], k);
        pb[k] = raw_sample<TIn, PX>(raw[1], k);
    }
#pragma unroll
    for (int t = 0; t < NMAX - 2; ++t) {
        if (t < n - 2) {
            const int pos = t < 2 ? 3 * t : 4 * t - 2;
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                pc[k] = raw_sample<TIn, PX>(raw[t + 2], k);
                const int a = pa[k], b = pb[k], c = pc[k];
                const int cur = a + b;
                uint32_t g = 0;
                if (t >= 2) g = shin_i(g, (t & 1) ? prev1[k] : prev0[k], cur);
                g = shin_i(g, a, avc[k]);
                g = shin_i(g, a, c);
                g = shin_i(g, a, b);
                put_group(w[k], pos, t >= 2 ? 4 : 3, g);
                if (t & 1) prev1[k] = cur; else prev0[k] = cur;
                pa[k] = b;
                pb[k] = c;
            }
        }
    }

    // closing bits: a<b, a<av, b<av, S(n-4) < S(n-2) (sentinel -1 for n < 4);
    // pa, pb now hold frames n - 2, n - 1
    const int P = n >= 4 ? 4 * n - 10 : 3 * (n - 2);
    const int wi = P >> 5, sh = P & 31;
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int a = pa[k], b = pb[k];
        const int prevf = n < 4 ? -1 : (((n - 2) & 1) ? prev1[k] : prev0[k]);
        uint32_t g = 0;
        g = shin_i(g, prevf, a + b);
        g = shin_i(g, b, avc[k]);
        g = shin_i(g, a, avc[k]);
        g = shin_i(g, a, b);
#pragma unroll
        for (int j = 0; j < WORDS; ++j) {
            if (j == wi) w[k][j] |= g << sh;
            if (j == wi + 1 && sh > 28) w[k][j] |= g >> (32 - sh);
        }
    }
    store_px<WORDS, PX>(out, W, row, col, w, out_words);
}

// full mode, compile-time N (descriptor_transform.hpp:75-123); the whole
// emission sequence is shifted in and each word bit-reversed at the end.
template<typename TIn, int N>
__global__ __launch_bounds__(256) void k_transform_full_v4(const Stack f0, const Stack f1,
                                                           int W, int out_words,
                                                           uint32_t* __restrict__ out0,
                                                           uint32_t* __restrict__ out1) {
    constexpr int BITS = N * N - 2 * N + 3;
    constexpr int WORDS = BITS <= 32 ? 1 : BITS <= 64 ? 2 : BITS <= 128 ? 4 : 8;
    const int col = (blockIdx.x * 256 + threadIdx.x) * TPX, row = blockIdx.y;
    if (col >= W)
        return;
    const Stack& f = blockIdx.z ? f1 : f0;
    uint32_t* __restrict__ out = blockIdx.z ? out1 : out0;

    float p[N][TPX];
    float av[TPX] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < N; ++t) {
        load_px4<TIn>(f, t, row, col, W, p[t]);
#pragma unroll
        for (int k = 0; k < TPX; ++k) av[k] += p[t][k];
    }
#pragma unroll
    for (int k = 0; k < TPX; ++k) av[k] /= (float)N;

    uint32_t w[TPX][WORDS];
#pragma unroll
    for (int k = 0; k < TPX; ++k) {
        uint32_t acc[WORDS];
#pragma unroll
        for (int j = 0; j < WORDS; ++j) acc[j] = 0u;
        int pos = 0;
        auto emit = [&](float x, float y) {
            acc[pos >> 5] = __builtin_amdgcn_alignbit(acc[pos >> 5], __float_as_uint(x - y), 31);
            ++pos;
        };
        float s[N - 1];
#pragma unroll
        for (int t = 0; t < N - 2; ++t) {
            emit(p[t][k], p[t + 1][k]);
            emit(p[t][k], p[t + 2][k]);
            emit(p[t][k], av[k]);
            s[t] = p[t][k] + p[t + 1][k];
        }
        s[N - 2] = p[N - 2][k] + p[N - 1][k];
        emit(p[N - 2][k], p[N - 1][k]);
        emit(p[N - 2][k], av[k]);
        emit(p[N - 1][k], av[k]);
#pragma unroll
        for (int t = 0; t < N - 1; ++t)
#pragma unroll
            for (int i = 0; i < N - 1; ++i) {
                if (i == t || i == t - 1 || i == t + 1)
                    continue;
                emit(s[t], s[i]);
            }
        // word j received min(32, BITS - 32 j) bits, first one highest
#pragma unroll
        for (int j = 0; j < WORDS; ++j) {
            const int m = BITS - 32 * j >= 32 ? 32 : (BITS - 32 * j > 0 ? BITS - 32 * j : 0);
            uint32_t r = __builtin_bitreverse32(acc[j]);
            w[k][j] = m == 0 ? 0u : (m == 32 ? r : r >> (32 - m));
        }
    }
    store_px(out, W, row, col, w, out_words);
}

}  // namespace bicos_hip
