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
__global__ __launch_bounds__(256) void k_transform_limited(const FrameTable f, int n, int W,
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
__global__ __launch_bounds__(256) void k_transform_full(const FrameTable f, int W, int out_words,
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

}  // namespace bicos_hip
