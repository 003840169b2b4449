// Learned disparity window: a small siamese CNN, its row correlation and
// its training, in hand-written HIP (no reference counterpart).
//
// Network (shared weights for both views), on a 1/4-resolution grid:
//   x0[t]  = (4x4 box mean of frame t) - (mean over the 4 frames), times in_scale
//   a1 = relu(conv3x3(x0, 4 -> 32)), a2 = relu(conv3x3(a1)), a3 = relu(conv3x3(a2)),
//   f  = conv3x3(a3)                       (32 features per cell)
//   C[k](Y, X) = <fL(Y, X), fR(Y, X - (D0 + k))> / sqrt(32),  k < K (runtime)
// The window of a pixel is centred on 4 * (D0 + k*) with k* the refined
// correlation peak of its cell. Training minimises the cross-entropy between
// softmax_k C and the plain BICOS disparity of the cell (split linearly over
// the two neighbouring bins). Weights and activations (x0, a1..a3, f) are
// BF16 values, every product is accumulated in FP32; FP32 master weights and
// Adam moments live in the caller's parameter buffers. The training kernels
// keep channel-planar [c][Y][X] FP32 tensors (holding BF16 values); the
// inference kernels keep cell-major [Y][X][c] BF16 tensors and run the
// 3x3x32 layers and the correlation on the matrix cores.
#pragma once

#include "device_common.hpp"

namespace bicos_hip {

constexpr int CNN_C = 32;        // channels
constexpr int CNN_IN = 4;        // input frames per view
constexpr int CNN_DS = 4;        // downsampling factor of the grid
constexpr int CNN_K = 80;        // default number of correlation bins (disparities -32 .. 284 px)
constexpr int CNN_D0 = -8;       // disparity (grid cells) of bin 0

// flat parameter layout: w1 b1 w2 b2 w3 b3 w4 b4, w[co][ci][3][3]
constexpr int CNN_W1 = CNN_C * CNN_IN * 9;
constexpr int CNN_WL = CNN_C * CNN_C * 9;
__host__ __device__ constexpr int cnn_w_off(int l) {
    return l == 0 ? 0 : CNN_W1 + CNN_C + (l - 1) * (CNN_WL + CNN_C);
}
__host__ __device__ constexpr int cnn_b_off(int l) { return cnn_w_off(l) + (l == 0 ? CNN_W1 : CNN_WL); }
constexpr int CNN_PARAMS = cnn_w_off(4);

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// round to nearest even BF16 (v_cvt_pk_bf16_f32), as float / as bits
__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }
__device__ __forceinline__ uint16_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

// x0 from the first CNN_IN frames of one view (strided stack)
template<typename TIn>
__global__ __launch_bounds__(256) void k_cnn_input(const Stack f, int W, int Hq, int Wq,
                                                   float in_scale, float* __restrict__ x0) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    float q[CNN_IN], m = 0.0f;
#pragma unroll
    for (int t = 0; t < CNN_IN; ++t) {
        float s = 0.0f;
#pragma unroll
        for (int dy = 0; dy < CNN_DS; ++dy) {
            const TIn* r = frame_row<TIn>(f, t, Y * CNN_DS + dy) + X * CNN_DS;
#pragma unroll
            for (int dx = 0; dx < CNN_DS; ++dx) s += (float)r[dx];
        }
        q[t] = s * (1.0f / (CNN_DS * CNN_DS));
        m += q[t];
    }
    m *= 1.0f / CNN_IN;
    const size_t plane = (size_t)Hq * Wq;
#pragma unroll
    for (int t = 0; t < CNN_IN; ++t) x0[t * plane + (size_t)Y * Wq + X] = bf16_round((q[t] - m) * in_scale);
}

// Inference weights as BF16 bits: layer 0 in [co][k] with k = tap * 4 + ci
// (36 values, zero-padded to CNN_K0 = 48: three MFMA steps of 16), layers
// 1..3 in [layer-1][co][tap][ci] so that one 16-byte load gives 8
// consecutive input channels of one tap.
constexpr int CNN_K0 = 48;
constexpr int CNN_MFMA_W0 = CNN_C * CNN_K0;
constexpr int CNN_MFMA_W = 3 * CNN_C * 9 * CNN_C;

__device__ __forceinline__ void cnn_pack_one(int i, const float* __restrict__ p, uint16_t* __restrict__ w0,
                                             uint16_t* __restrict__ wm) {
    if (i < CNN_MFMA_W0) {
        const int co = i / CNN_K0, k = i % CNN_K0, tap = k / CNN_IN, ci = k % CNN_IN;
        w0[i] = tap < 9 ? bf16_bits(p[cnn_w_off(0) + (co * CNN_IN + ci) * 9 + tap]) : (uint16_t)0;
    }
    if (i < CNN_MFMA_W) {
        const int l = i / (CNN_C * 9 * CNN_C), r = i % (CNN_C * 9 * CNN_C);
        const int co = r / (9 * CNN_C), tap = (r / CNN_C) % 9, ci = r % CNN_C;
        wm[i] = bf16_bits(p[cnn_w_off(l + 1) + (co * CNN_C + ci) * 9 + tap]);
    }
}

__global__ void k_cnn_pack_mfma(const float* __restrict__ p, uint16_t* __restrict__ w0,
                                uint16_t* __restrict__ wm) {
    cnn_pack_one(blockIdx.x * 256 + threadIdx.x, p, w0, wm);
}

// k_cnn_input for both views at once (view = blockIdx.z), written
// cell-major as BF16 bits [Y][X][CNN_IN]; VEC: the 4 samples of a cell row
// come in one load (stack base, frame stride and row step all multiples of
// 4 samples). Same arithmetic as k_cnn_input. The blocks of z = 2 pack the
// inference weights (k_cnn_pack_mfma's work) and clear zero[0 .. nzero)
// instead (a caller's counters, e.g. the fallback lists').
template<typename TIn, bool VEC>
__global__ __launch_bounds__(256) void k_cnn_input_pair(const Stack fa, const Stack fb, int Hq,
                                                        int Wq, float in_scale,
                                                        uint16_t* __restrict__ x0a,
                                                        uint16_t* __restrict__ x0b,
                                                        const float* __restrict__ params,
                                                        uint16_t* __restrict__ w0,
                                                        uint16_t* __restrict__ wm,
                                                        uint32_t* __restrict__ zero, int nzero) {
    static_assert(CNN_DS == 4, "one 4-sample load per cell row");
    if (blockIdx.z == 2) {
        const int threads = gridDim.x * gridDim.y * 256;
        const int first = (blockIdx.y * gridDim.x + blockIdx.x) * 256 + threadIdx.x;
        for (int i = first; i < CNN_MFMA_W; i += threads) cnn_pack_one(i, params, w0, wm);
        for (int i = first; i < nzero; i += threads) zero[i] = 0u;
        return;
    }
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const Stack& f = blockIdx.z ? fb : fa;
    float q[CNN_IN], m = 0.0f;
#pragma unroll
    for (int t = 0; t < CNN_IN; ++t) {
        float s = 0.0f;
#pragma unroll
        for (int dy = 0; dy < CNN_DS; ++dy) {
            const TIn* r = frame_row<TIn>(f, t, Y * CNN_DS + dy) + X * CNN_DS;
            TIn v[4];
            if (VEC) {
                if (sizeof(TIn) == 2) {
                    const uint2 w = *(const uint2*)r;
                    v[0] = (TIn)(w.x & 0xFFFF); v[1] = (TIn)(w.x >> 16);
                    v[2] = (TIn)(w.y & 0xFFFF); v[3] = (TIn)(w.y >> 16);
                } else {
                    const uint32_t w = *(const uint32_t*)r;
                    v[0] = (TIn)(w & 0xFF); v[1] = (TIn)((w >> 8) & 0xFF);
                    v[2] = (TIn)((w >> 16) & 0xFF); v[3] = (TIn)(w >> 24);
                }
            } else {
#pragma unroll
                for (int dx = 0; dx < 4; ++dx) v[dx] = r[dx];
            }
#pragma unroll
            for (int dx = 0; dx < CNN_DS; ++dx) s += (float)v[dx];
        }
        q[t] = s * (1.0f / (CNN_DS * CNN_DS));
        m += q[t];
    }
    m *= 1.0f / CNN_IN;
    static_assert(CNN_IN == 4, "one 8-byte store per cell");
    uint16_t h[CNN_IN];
#pragma unroll
    for (int t = 0; t < CNN_IN; ++t) h[t] = bf16_bits((q[t] - m) * in_scale);
    uint16_t* x0 = blockIdx.z ? x0b : x0a;
    *(uint2*)(x0 + ((size_t)Y * Wq + X) * CNN_IN) =
        make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
}

// BF16 copy of the weights (biases stay FP32)
__global__ void k_cnn_bf16_weights(const float* __restrict__ p, float* __restrict__ q) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= CNN_PARAMS) return;
    bool is_bias = false;
#pragma unroll
    for (int l = 0; l < 4; ++l)
        if (i >= cnn_b_off(l) && i < cnn_b_off(l) + CNN_C) is_bias = true;
    q[i] = is_bias ? p[i] : bf16_round(p[i]);
}

// 3x3 convolution, zero padding; one lane per output cell computes all 32
// output channels. Weights (already BF16-valued) staged in LDS.
template<int CIN, bool RELU>
__global__ __launch_bounds__(256) void k_cnn_conv(const float* __restrict__ x, const float* __restrict__ w,
                                                  const float* __restrict__ b, int Hq, int Wq,
                                                  float* __restrict__ y) {
    __shared__ float ws[CNN_C * CIN * 9];
    for (int i = threadIdx.x; i < CNN_C * CIN * 9; i += 256) ws[i] = w[i];
    __syncthreads();
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq;
    float acc[CNN_C];
#pragma unroll
    for (int co = 0; co < CNN_C; ++co) acc[co] = b[co];
    for (int ci = 0; ci < CIN; ++ci) {
        float v[9];
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int yy = Y + ky - 1, xx = X + kx - 1;
                v[ky * 3 + kx] = (yy >= 0 && yy < Hq && xx >= 0 && xx < Wq)
                                     ? x[ci * plane + (size_t)yy * Wq + xx] : 0.0f;
            }
#pragma unroll
        for (int co = 0; co < CNN_C; ++co)
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[co] = __builtin_fmaf(ws[(co * CIN + ci) * 9 + k], v[k], acc[co]);
    }
#pragma unroll
    for (int co = 0; co < CNN_C; ++co)
        y[co * plane + (size_t)Y * Wq + X] = bf16_round(RELU ? fmaxf(acc[co], 0.0f) : acc[co]);
}

// correlation volume C[k][Y][X] (masked bins: -1e30)
__global__ __launch_bounds__(256) void k_cnn_corr(const float* __restrict__ fl, const float* __restrict__ fr,
                                                  int Hq, int Wq, int K, float* __restrict__ C) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq, at = (size_t)Y * Wq + X;
    float l[CNN_C];
#pragma unroll
    for (int c = 0; c < CNN_C; ++c) l[c] = fl[c * plane + at];
    for (int k = 0; k < K; ++k) {
        const int xr = X - (CNN_D0 + k);
        float s = -1e30f;
        if (xr >= 0 && xr < Wq) {
            s = 0.0f;
#pragma unroll
            for (int c = 0; c < CNN_C; ++c) s = __builtin_fmaf(l[c], fr[c * plane + (size_t)Y * Wq + xr], s);
            s *= 0.17677669529663687f;  // 1 / sqrt(32)
        }
        C[k * plane + at] = s;
    }
}

// softmax cross-entropy against the label disparity (full-resolution px,
// NaN = unlabelled); writes dC (scaled by 1/count later) and the cell loss
__global__ __launch_bounds__(256) void k_cnn_softmax_ce(const float* __restrict__ C,
                                                        const float* __restrict__ label, int Hq,
                                                        int Wq, int K, float* __restrict__ dC,
                                                        float* __restrict__ cell_loss) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq, at = (size_t)Y * Wq + X;
    const float lab = label[at];
    float mx = -1e30f;
    for (int k = 0; k < K; ++k) mx = fmaxf(mx, C[k * plane + at]);
    float z = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float c = C[k * plane + at];
        z += c > -1e29f ? __expf(c - mx) : 0.0f;
    }
    const float lz = mx + __logf(z);
    float loss = 0.0f;
    const bool has = !__builtin_isnan(lab);
    float pos = has ? lab / CNN_DS - CNN_D0 : 0.0f;   // fractional bin
    bool in = has && pos >= 0.0f && pos <= (float)(K - 1);
    const int k0 = in ? (int)floorf(pos) : -1;
    const float w1 = in ? pos - k0 : 0.0f;
    // a cell whose target bins fall outside the right image is unlabelled
    const int xr0 = X - (CNN_D0 + k0);
    in = in && xr0 >= 0 && xr0 < Wq && (w1 == 0.0f || xr0 - 1 >= 0);
    for (int k = 0; k < K; ++k) {
        const float c = C[k * plane + at];
        const float p = c > -1e29f ? __expf(c - lz) : 0.0f;
        const float tgt = !in ? 0.0f : (k == k0 ? 1.0f - w1 : (k == k0 + 1 ? w1 : 0.0f));
        if (in && tgt > 0.0f) loss -= tgt * (c - lz);
        dC[k * plane + at] = in ? p - tgt : 0.0f;
    }
    cell_loss[at] = in ? loss : 0.0f;
    cell_loss[plane + at] = in ? 1.0f : 0.0f;
}

// gradients of the correlation w.r.t. both feature maps
__global__ __launch_bounds__(256) void k_cnn_corr_bwd(const float* __restrict__ fl, const float* __restrict__ fr,
                                                      const float* __restrict__ dC, float scale, int Hq,
                                                      int Wq, int K, float* __restrict__ dfl,
                                                      float* __restrict__ dfr) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq, at = (size_t)Y * Wq + X;
    const float s = scale * 0.17677669529663687f;
    float gl[CNN_C], gr[CNN_C];
#pragma unroll
    for (int c = 0; c < CNN_C; ++c) gl[c] = gr[c] = 0.0f;
    for (int k = 0; k < K; ++k) {
        const int d = CNN_D0 + k;
        const int xr = X - d;            // this cell as the left one
        if (xr >= 0 && xr < Wq) {
            const float g = dC[k * plane + at] * s;
            if (g != 0.0f)
#pragma unroll
                for (int c = 0; c < CNN_C; ++c) gl[c] = __builtin_fmaf(g, fr[c * plane + (size_t)Y * Wq + xr], gl[c]);
        }
        const int xl = X + d;            // this cell as the right one
        if (xl >= 0 && xl < Wq) {
            const float g = dC[k * plane + (size_t)Y * Wq + xl] * s;
            if (g != 0.0f)
#pragma unroll
                for (int c = 0; c < CNN_C; ++c) gr[c] = __builtin_fmaf(g, fl[c * plane + (size_t)Y * Wq + xl], gr[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CNN_C; ++c) {
        dfl[c * plane + at] = gl[c];
        dfr[c * plane + at] = gr[c];
    }
}

// dy_pre = dy * (y > 0) for layers followed by a ReLU (in place on dy)
__global__ void k_cnn_relu_bwd(const float* __restrict__ y, float* __restrict__ dy, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count && !(y[i] > 0.0f)) dy[i] = 0.0f;
}

// dx[ci] = sum_co sum_k w[co][ci][k] * dy[co] at the mirrored offset
template<int CIN>
__global__ __launch_bounds__(256) void k_cnn_conv_bwd_data(const float* __restrict__ dy, const float* __restrict__ w,
                                                           int Hq, int Wq, float* __restrict__ dx) {
    __shared__ float ws[CNN_C * CIN * 9];
    for (int i = threadIdx.x; i < CNN_C * CIN * 9; i += 256) ws[i] = w[i];
    __syncthreads();
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq;
    float acc[CIN];
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.0f;
    for (int co = 0; co < CNN_C; ++co) {
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                // y(Y', X') used x(Y' + ky - 1, X' + kx - 1): here Y' = Y - ky + 1
                const int yy = Y - ky + 1, xx = X - kx + 1;
                if (yy < 0 || yy >= Hq || xx < 0 || xx >= Wq) continue;
                const float g = dy[co * plane + (size_t)yy * Wq + xx];
#pragma unroll
                for (int ci = 0; ci < CIN; ++ci) acc[ci] = __builtin_fmaf(ws[(co * CIN + ci) * 9 + ky * 3 + kx], g, acc[ci]);
            }
    }
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) dx[ci * plane + (size_t)Y * Wq + X] = acc[ci];
}

// per-block partial weight/bias gradients: one block per (co, row band);
// lane = tap-input pair, summed over the band's cells (deterministic order)
template<int CIN>
__global__ __launch_bounds__(256) void k_cnn_conv_bwd_weight(const float* __restrict__ x, const float* __restrict__ dy,
                                                             int Hq, int Wq, int rows_per_band,
                                                             float* __restrict__ partial) {
    constexpr int NT = CIN * 9 + 1;   // weights of one output channel + its bias
    const int co = blockIdx.x, band = blockIdx.y;
    const int y0 = band * rows_per_band, y1 = min(Hq, y0 + rows_per_band);
    const size_t plane = (size_t)Hq * Wq;
    for (int j = threadIdx.x; j < NT; j += 256) {
        float s = 0.0f;
        if (j == NT - 1) {
            for (int Y = y0; Y < y1; ++Y)
                for (int X = 0; X < Wq; ++X) s += dy[co * plane + (size_t)Y * Wq + X];
        } else {
            const int ci = j / 9, ky = (j % 9) / 3, kx = j % 3;
            for (int Y = y0; Y < y1; ++Y) {
                const int yy = Y + ky - 1;
                if (yy < 0 || yy >= Hq) continue;
                for (int X = 0; X < Wq; ++X) {
                    const int xx = X + kx - 1;
                    if (xx < 0 || xx >= Wq) continue;
                    s = __builtin_fmaf(dy[co * plane + (size_t)Y * Wq + X], x[ci * plane + (size_t)yy * Wq + xx], s);
                }
            }
        }
        partial[((size_t)band * CNN_C + co) * NT + j] = s;
    }
}

// sums band partials into the flat gradient (accumulating if acc != 0)
template<int CIN>
__global__ void k_cnn_reduce_weight(const float* __restrict__ partial, int bands, int layer,
                                    int acc, float* __restrict__ grad) {
    constexpr int NT = CIN * 9 + 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= CNN_C * NT) return;
    const int co = i / NT, j = i % NT;
    float s = 0.0f;
    for (int b = 0; b < bands; ++b) s += partial[((size_t)b * CNN_C + co) * NT + j];
    const int dst = j == NT - 1 ? cnn_b_off(layer) + co : cnn_w_off(layer) + co * CIN * 9 + j;
    grad[dst] = acc ? grad[dst] + s : s;
}

// Adam on the flat parameters; the gradient is grad / max(count, 1)
__global__ void k_cnn_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, const float* __restrict__ count, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= CNN_PARAMS) return;
    const float scale = 1.0f / fmaxf(count ? count[0] : 1.0f, 1.0f);
    const float gi = g[i] * scale;
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// window of every cell from its correlation peak: centre 4 (D0 + k*) px
// with k* the argmax bin refined by a parabola through its neighbours,
// half-width hw_lo where the peak's softmax probability is at least p_conf,
// else hw_hi
__global__ __launch_bounds__(256) void k_cnn_window(const float* __restrict__ C, int Hq, int Wq, int K,
                                                    float p_conf, int hw_lo, int hw_hi,
                                                    float* __restrict__ center,
                                                    int16_t* __restrict__ halfwidth,
                                                    float* __restrict__ peak_prob) {
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= Wq)
        return;
    const size_t plane = (size_t)Hq * Wq, at = (size_t)Y * Wq + X;
    int kb = 0;
    float mx = -1e30f;
    for (int k = 0; k < K; ++k) {
        const float c = C[k * plane + at];
        if (c > mx) { mx = c; kb = k; }
    }
    float z = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float c = C[k * plane + at];
        z += c > -1e29f ? __expf(c - mx) : 0.0f;
    }
    const float pk = 1.0f / z;
    float kf = (float)kb;
    if (kb > 0 && kb < K - 1) {
        const float cm = C[(kb - 1) * plane + at], cp = C[(kb + 1) * plane + at];
        if (cm > -1e29f && cp > -1e29f) {
            const float den = cm - 2.0f * mx + cp;
            if (den < 0.0f) kf += fminf(fmaxf(0.5f * (cm - cp) / den, -0.5f), 0.5f);
        }
    }
    center[at] = CNN_DS * (CNN_D0 + kf);
    halfwidth[at] = (int16_t)(pk >= p_conf ? hw_lo : hw_hi);
    if (peak_prob) peak_prob[at] = pk;
}

/* ------------------------------------------------------------ inference */

// layer 0 (4 -> 32 channels, ReLU) on v_mfma_f32_32x32x16_bf16, from the
// cell-major x0 of view blockIdx.z: one wave per 32-cell tile of a row,
// D[co][cell] = W0[co][k] x In[k][cell] over k = tap * 4 + ci (three steps;
// lane half g of step s holds taps 4 s + 2 g and 4 s + 2 g + 1). Output as
// k_cnn_conv_mfma.
__global__ __launch_bounds__(256) void k_cnn_conv0_mfma(const uint16_t* __restrict__ xa,
                                                        const uint16_t* __restrict__ xb,
                                                        const uint16_t* __restrict__ w0,
                                                        const float* __restrict__ bias, int Hq,
                                                        int Wq, uint16_t* __restrict__ ya,
                                                        uint16_t* __restrict__ yb) {
    const int lane = threadIdx.x & 63, nl = lane & 31, g = lane >> 5;
    const int tiles_row = (Wq + 31) / 32;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int Y = wave / tiles_row, X = (wave % tiles_row) * 32 + nl;
    if (Y >= Hq)
        return;
    const uint16_t* x = blockIdx.z ? xb : xa;
    uint16_t* y = blockIdx.z ? yb : ya;
    const auto tap_cell = [&](int tap) {  // 4 channels of the tap's cell, zero outside
        uint2 v = make_uint2(0, 0);
        const int yy = Y + tap / 3 - 1, xx = X + tap % 3 - 1;
        if (tap < 9 && yy >= 0 && yy < Hq && xx >= 0 && xx < Wq)
            v = *(const uint2*)(x + ((size_t)yy * Wq + xx) * CNN_IN);
        return v;
    };
    f32x16 acc = {0};
#pragma unroll
    for (int st = 0; st < CNN_K0 / 16; ++st) {
        const bf16x8 wa = __builtin_bit_cast(bf16x8, *(const uint4*)(w0 + nl * CNN_K0 + 16 * st + 8 * g));
        const uint2 lo = tap_cell(4 * st + 2 * g), hi = tap_cell(4 * st + 2 * g + 1);
        const bf16x8 bx = __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, bx, acc, 0, 0, 0);
    }
    if (X < Wq) {
        uint16_t* dst = y + ((size_t)Y * Wq + X) * CNN_C + 4 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                o[r] = fmaxf(acc[4 * q + r] + bias[8 * q + 4 * g + r], 0.0f);
            *(uint2*)(dst + 8 * q) = make_uint2((uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16),
                                                (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16));
        }
    }
}

// 16-byte load of 8 BF16 channels, zero outside the grid
__device__ __forceinline__ bf16x8 load_ch8(const uint16_t* t, int Hq, int Wq, int Y, int X, int c0) {
    uint4 v = make_uint4(0, 0, 0, 0);
    if (Y >= 0 && Y < Hq && X >= 0 && X < Wq)
        v = *(const uint4*)(t + ((size_t)Y * Wq + X) * CNN_C + c0);
    return __builtin_bit_cast(bf16x8, v);
}

// 3x3 conv, 32 -> 32 channels, on v_mfma_f32_32x32x16_bf16: D[co][cell] =
// W[co][k] x In[k][cell] with k = (tap, ci) (K = 288 = 18 steps of 16). A
// workgroup computes CNN_TROWS rows of one 32-cell tile column, a wave per
// row: the input patch (rows and columns one beyond on each side, zero
// outside the grid) is staged in LDS in one round of loads, issued together
// with the wave's weight operands, and the B operands are read from there.
// Output lane layout: lane l, accumulator i -> cell l % 32, channel 8 (i / 4)
// + 4 (l / 32) + i % 4, stored as 4 runs of 4 channels. View = blockIdx.z.
constexpr int CNN_TROWS = 4;
constexpr int CNN_CONV_THREADS = CNN_TROWS * 64;
constexpr int CNN_PCOLS = 34;    // patch columns
constexpr int CNN_PSTRIDE = 40;  // patch cell stride in BF16 values (64 B of channels, 16 B pad)

template<bool RELU>
__global__ __launch_bounds__(CNN_CONV_THREADS) void k_cnn_conv_mfma(const uint16_t* __restrict__ xa,
                                                       const uint16_t* __restrict__ xb,
                                                       const uint16_t* __restrict__ wm,
                                                       const float* __restrict__ bias, int Hq,
                                                       int Wq, uint16_t* __restrict__ ya,
                                                       uint16_t* __restrict__ yb) {
    constexpr int PROWS = CNN_TROWS + 2, CHUNKS = PROWS * CNN_PCOLS * 4;  // 16-byte chunks
    __shared__ uint4 patch_raw[PROWS * CNN_PCOLS * CNN_PSTRIDE / 8];
    uint16_t* patch = (uint16_t*)patch_raw;
    const int lane = threadIdx.x & 63, nl = lane & 31, g = lane >> 5, wv = threadIdx.x >> 6;
    const int tiles_row = (Wq + 31) / 32;
    const int X0 = (blockIdx.x % tiles_row) * 32, Y0 = (blockIdx.x / tiles_row) * CNN_TROWS;
    const uint16_t* x = blockIdx.z ? xb : xa;
    uint16_t* y = blockIdx.z ? yb : ya;
    bf16x8 wa[18];
#pragma unroll
    for (int s = 0; s < 18; ++s)
        wa[s] = __builtin_bit_cast(bf16x8, *(const uint4*)(wm + ((size_t)nl * 9 + (s >> 1)) * CNN_C + 16 * (s & 1) + 8 * g));
#pragma unroll
    for (int q = 0; q < (CHUNKS + CNN_CONV_THREADS - 1) / CNN_CONV_THREADS; ++q) {
        const int i = q * CNN_CONV_THREADS + threadIdx.x;
        if (i < CHUNKS) {
            const int cell = i >> 2, part = i & 3, pr = cell / CNN_PCOLS, pc = cell % CNN_PCOLS;
            const int yy = Y0 - 1 + pr, xx = X0 - 1 + pc;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < Hq && xx >= 0 && xx < Wq)
                v = *(const uint4*)(x + ((size_t)yy * Wq + xx) * CNN_C + 8 * part);
            *(uint4*)(patch + cell * CNN_PSTRIDE + 8 * part) = v;
        }
    }
    float bv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) bv[i] = bias[8 * (i / 4) + 4 * g + (i % 4)];
    __syncthreads();
    const int Y = Y0 + wv, X = X0 + nl;
    if (Y >= Hq)
        return;
    f32x16 acc = {0};
#pragma unroll
    for (int s = 0; s < 18; ++s) {
        const int tap = s >> 1, ky = tap / 3, kx = tap % 3;
        const bf16x8 bx = *(const bf16x8*)(patch + ((wv + ky) * CNN_PCOLS + nl + kx) * CNN_PSTRIDE + 16 * (s & 1) + 8 * g);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], bx, acc, 0, 0, 0);
    }
    if (X < Wq) {
        uint16_t* dst = y + ((size_t)Y * Wq + X) * CNN_C + 4 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float v = acc[4 * q + r] + bv[4 * q + r];
                o[r] = RELU ? fmaxf(v, 0.0f) : v;
            }
            *(uint2*)(dst + 8 * q) = make_uint2((uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16),
                                                (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16));
        }
    }
}

// The four layers in one kernel: a workgroup of 8 waves computes the
// features of a CNN_FT_H x CNN_FT_W tile of cells. The tile's input (4-cell
// halo) is staged in LDS once, and each layer's activations over the region
// the later layers need (1 cell less per side per layer) stay in LDS: a1 in
// buffer A, a2 in B, a3 in A again; only the features go to memory. Each
// region row is one 32-cell MFMA row tile (the regions are at most 32
// cells wide; lanes past a region's width compute values that are not
// stored). Every output takes the same products in the same order as
// k_cnn_conv0_mfma / k_cnn_conv_mfma, with the same bias, ReLU and BF16
// rounding, and cells outside the grid are stored as zero (the layers' zero
// padding), so the features equal the per-layer kernels' bit for bit.
constexpr int CNN_FT_W = 26, CNN_FT_H = 8;  // feature cells per workgroup
constexpr int CNN_FT_COLS = 34;             // stored cells per region row (32 + 2 read past)
constexpr int CNN_FT_STRIDE = 40;           // BF16 values per stored cell (32 channels, 16 B pad)
constexpr int CNN_FT_THREADS = 512;

// The 16-byte chunks of one 3x3x32 layer's weights (wm layout [co][tap][ci])
// that a thread of the workgroup stages at the start of their layer. Chunk
// c = co * 36 + 4 tap + 2 h + g (channels 16 h + 8 g .. + 7 of the tap) is
// the A operand of lane co of lane half g at K step s = 2 tap + h; it is
// staged at (c % 36) * 32 + co = (2 s + g) * 32 + co, so that the 64 lanes of
// each step read 64 consecutive chunks (no LDS bank conflicts).
constexpr int CNN_WQ = CNN_C * 9 * CNN_C / 8;  // chunks per layer
static_assert(CNN_WQ > 2 * CNN_FT_THREADS && CNN_WQ <= 3 * CNN_FT_THREADS, "three chunks per thread");
struct CnnWeightChunks {  // this thread's chunks tid, tid + 512, tid + 1024 (the last if < CNN_WQ)
    uint4 c0, c1, c2;
};
__device__ __forceinline__ void cnn_ft_weights_load(const uint16_t* __restrict__ wl, CnnWeightChunks& w) {
    // unconditional (clamped) loads, so that nothing waits for them where they are issued
    const uint4* q = (const uint4*)wl;
    w.c0 = q[threadIdx.x];
    w.c1 = q[threadIdx.x + CNN_FT_THREADS];
    w.c2 = q[min((int)threadIdx.x + 2 * CNN_FT_THREADS, CNN_WQ - 1)];
}
__device__ __forceinline__ void cnn_ft_weights_stage(const CnnWeightChunks& w, uint4* __restrict__ wtmp) {
    const auto put = [&](int c, const uint4& v) { wtmp[(c % 36) * 32 + c / 36] = v; };
    put(threadIdx.x, w.c0);
    put(threadIdx.x + CNN_FT_THREADS, w.c1);
    if ((int)threadIdx.x + 2 * CNN_FT_THREADS < CNN_WQ) put(threadIdx.x + 2 * CNN_FT_THREADS, w.c2);
}

// one 3x3x32 layer over the stored region `in` (origin one cell up-left of
// the output region's), rows / width of the output region; TO_LDS: into the
// stored region `out`, else into the in-grid cells of the feature map y.
// The layer's weights wl are staged through `wtmp` (LDS the layer does not
// read and has not yet written) once per workgroup, and each wave takes its
// operands from there (holding the next layer's chunks in registers during
// the layer would cost the kernel a wave per SIMD).
template<bool RELU, bool TO_LDS>
__device__ __forceinline__ void cnn_ft_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                             int rows, int width, int oy, int ox, int Hq, int Wq,
                                             const uint16_t* __restrict__ wl,
                                             const float* __restrict__ bias, uint16_t* __restrict__ y,
                                             uint4* __restrict__ wtmp) {
    const int lane = threadIdx.x & 63, nl = lane & 31, g = lane >> 5, wave = threadIdx.x >> 6;
    {
        CnnWeightChunks w;
        cnn_ft_weights_load(wl, w);
        cnn_ft_weights_stage(w, wtmp);
    }
    __syncthreads();
    bf16x8 wa[18];
#pragma unroll
    for (int s = 0; s < 18; ++s)  // chunk nl * 36 + 4 tap + 2 (s & 1) + g, tap = s >> 1
        wa[s] = __builtin_bit_cast(bf16x8, wtmp[(2 * s + g) * 32 + nl]);
    __syncthreads();
    float bv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) bv[i] = bias[8 * (i / 4) + 4 * g + (i % 4)];
    for (int r = wave; r < rows; r += CNN_FT_THREADS / 64) {
        f32x16 acc = {0};
#pragma unroll
        for (int s = 0; s < 18; ++s) {
            const int tap = s >> 1, ky = tap / 3, kx = tap % 3;
            const bf16x8 bx = *(const bf16x8*)(in + ((r + ky) * CNN_FT_COLS + nl + kx) * CNN_FT_STRIDE + 16 * (s & 1) + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], bx, acc, 0, 0, 0);
        }
        const int Y = oy + r, X = ox + nl;
        const bool inside = Y >= 0 && Y < Hq && X >= 0 && X < Wq;
        uint32_t pk[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float o[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float v = acc[4 * q + k] + bv[4 * q + k];
                o[k] = inside ? (RELU ? fmaxf(v, 0.0f) : v) : 0.0f;
            }
            pk[2 * q] = (uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16);
            pk[2 * q + 1] = (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16);
        }
        if constexpr (TO_LDS) {
            if (nl < width) {
                uint16_t* dst = out + (r * CNN_FT_COLS + nl) * CNN_FT_STRIDE + 4 * g;
#pragma unroll
                for (int q = 0; q < 4; ++q) *(uint2*)(dst + 8 * q) = make_uint2(pk[2 * q], pk[2 * q + 1]);
            }
        } else {
            if (nl < width && inside) {
                uint16_t* dst = y + ((size_t)Y * Wq + X) * CNN_C + 4 * g;
#pragma unroll
                for (int q = 0; q < 4; ++q) *(uint2*)(dst + 8 * q) = make_uint2(pk[2 * q], pk[2 * q + 1]);
            }
        }
    }
}

__global__ __launch_bounds__(CNN_FT_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_cnn_features_fused(
    const uint16_t* __restrict__ xa, const uint16_t* __restrict__ xb, const uint16_t* __restrict__ w0,
    const uint16_t* __restrict__ wm, const float* __restrict__ params, int Hq, int Wq,
    uint16_t* __restrict__ ya, uint16_t* __restrict__ yb) {
    constexpr int XR = CNN_FT_H + 8, A_ROWS = CNN_FT_H + 6, B_ROWS = CNN_FT_H + 4;
    __shared__ uint2 xin[XR * CNN_FT_COLS];  // input cells (4 channels), rows Y0-4.., cols X0-4..
    __shared__ uint4 bufA_raw[A_ROWS * CNN_FT_COLS * CNN_FT_STRIDE / 8];
    __shared__ uint4 bufB_raw[B_ROWS * CNN_FT_COLS * CNN_FT_STRIDE / 8];
    static_assert(B_ROWS * CNN_FT_COLS * CNN_FT_STRIDE >= CNN_C * 9 * CNN_C, "a layer's weights fit either buffer");
    uint16_t* bufA = (uint16_t*)bufA_raw;
    uint16_t* bufB = (uint16_t*)bufB_raw;
    const int lane = threadIdx.x & 63, nl = lane & 31, g = lane >> 5, wave = threadIdx.x >> 6;
    const int tiles_x = (Wq + CNN_FT_W - 1) / CNN_FT_W;
    const int X0 = (blockIdx.x % tiles_x) * CNN_FT_W, Y0 = (blockIdx.x / tiles_x) * CNN_FT_H;
    const uint16_t* x = blockIdx.z ? xb : xa;
    uint16_t* y = blockIdx.z ? yb : ya;
    const size_t wl = (size_t)CNN_C * 9 * CNN_C;
    // all loads in flight together: clamped addresses, zero selected outside the grid
    constexpr int XIN_ITERS = (XR * CNN_FT_COLS + CNN_FT_THREADS - 1) / CNN_FT_THREADS;
    uint2 xv[XIN_ITERS];
#pragma unroll
    for (int q = 0; q < XIN_ITERS; ++q) {
        const int i = threadIdx.x + q * CNN_FT_THREADS;
        const int yy = Y0 - 4 + i / CNN_FT_COLS, xx = X0 - 4 + i % CNN_FT_COLS;
        const uint2 v = *(const uint2*)(x + ((size_t)min(max(yy, 0), Hq - 1) * Wq + min(max(xx, 0), Wq - 1)) * CNN_IN);
        xv[q] = (yy >= 0 && yy < Hq && xx >= 0 && xx < Wq) ? v : make_uint2(0, 0);
    }
#pragma unroll
    for (int q = 0; q < XIN_ITERS; ++q) {
        const int i = threadIdx.x + q * CNN_FT_THREADS;
        if (i < XR * CNN_FT_COLS) xin[i] = xv[q];
    }
    __syncthreads();
    {   // layer 0 (4 -> 32, ReLU) into A: rows Y0-3.., cols X0-3.. (k_cnn_conv0_mfma's products)
        bf16x8 wa[CNN_K0 / 16];
#pragma unroll
        for (int st = 0; st < CNN_K0 / 16; ++st)
            wa[st] = __builtin_bit_cast(bf16x8, *(const uint4*)(w0 + nl * CNN_K0 + 16 * st + 8 * g));
        float bv[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 b4 = *(const float4*)(params + cnn_b_off(0) + 8 * q + 4 * g);
            bv[4 * q] = b4.x; bv[4 * q + 1] = b4.y; bv[4 * q + 2] = b4.z; bv[4 * q + 3] = b4.w;
        }
        for (int r = wave; r < A_ROWS; r += CNN_FT_THREADS / 64) {
            f32x16 acc = {0};
#pragma unroll
            for (int st = 0; st < CNN_K0 / 16; ++st) {
                const int t0 = 4 * st + 2 * g, t1 = t0 + 1;  // 4 channels of each tap's cell
                uint2 lo = make_uint2(0, 0), hi = make_uint2(0, 0);
                if (t0 < 9) lo = xin[(r + t0 / 3) * CNN_FT_COLS + nl + t0 % 3];
                if (t1 < 9) hi = xin[(r + t1 / 3) * CNN_FT_COLS + nl + t1 % 3];
                const bf16x8 bx = __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[st], bx, acc, 0, 0, 0);
            }
            const int Y = Y0 - 3 + r, X = X0 - 3 + nl;
            const bool inside = Y >= 0 && Y < Hq && X >= 0 && X < Wq;
            uint16_t* dst = bufA + (r * CNN_FT_COLS + nl) * CNN_FT_STRIDE + 4 * g;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float o[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    o[k] = inside ? fmaxf(acc[4 * q + k] + bv[4 * q + k], 0.0f) : 0.0f;
                *(uint2*)(dst + 8 * q) = make_uint2((uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16),
                                                    (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16));
            }
        }
    }
    __syncthreads();
    // weights staged in the buffer each layer writes (B, A) or no longer
    // needs (B, after a3)
    cnn_ft_layer<true, true>(bufA, bufB, B_ROWS, CNN_FT_W + 4, Y0 - 2, X0 - 2, Hq, Wq, wm,
                             params + cnn_b_off(1), nullptr, bufB_raw);
    __syncthreads();
    cnn_ft_layer<true, true>(bufB, bufA, CNN_FT_H + 2, CNN_FT_W + 2, Y0 - 1, X0 - 1, Hq, Wq, wm + wl,
                             params + cnn_b_off(2), nullptr, bufA_raw);
    __syncthreads();
    cnn_ft_layer<false, false>(bufA, nullptr, CNN_FT_H, CNN_FT_W, Y0, X0, Hq, Wq, wm + 2 * wl,
                               params + cnn_b_off(3), y, bufB_raw);
}

// Correlation and window of every cell of a 32-cell tile (one wave): the
// scores <fL(X), fR(X')> of the tile's cells against the right cells
// X' = X - (D0 + k) come from v_mfma_f32_32x32x16_bf16 over right tiles of
// 32 cells, land in LDS as s[cell][k], and lanes 0..31 then take the
// argmax, the softmax peak probability and the refined centre exactly as
// k_cnn_window does. K <= CNN_KMAX_MFMA.
constexpr int CNN_KMAX_MFMA = 160;

// dynamic LDS of k_cnn_corr_window for K bins
inline size_t cnn_corr_window_lds(int K) { return (size_t)4 * (32 * (K + 1) + 1) * sizeof(float); }

__global__ __launch_bounds__(256) void k_cnn_corr_window(const uint16_t* __restrict__ fl,
                                                         const uint16_t* __restrict__ fr, int Hq,
                                                         int Wq, int K, float p_conf, int hw_lo,
                                                         int hw_hi, float* __restrict__ center,
                                                         int16_t* __restrict__ halfwidth,
                                                         float* __restrict__ peak_prob) {
    extern __shared__ float sc_dyn[];
    const int lane = threadIdx.x & 63, nl = lane & 31, g = lane >> 5, wv = threadIdx.x >> 6;
    const int tiles_row = (Wq + 31) / 32;
    const int wave = blockIdx.x * 4 + wv;
    const int Y = wave / tiles_row, X0 = (wave % tiles_row) * 32;
    if (Y >= Hq)
        return;
    const int ld = K + 1;
    float* S = sc_dyn + (size_t)wv * (32 * ld + 1);  // S[cell * ld + k], then a spare slot
    // left operands: cell X0 + nl, channels 16 h + 8 g .. +7
    const bf16x8 a0 = load_ch8(fl, Hq, Wq, Y, X0 + nl, 8 * g);
    const bf16x8 a1 = load_ch8(fl, Hq, Wq, Y, X0 + nl, 16 + 8 * g);
    // right cells span [X0 - D0 - K + 1, X0 + 31 - D0]
    const int rb = X0 - CNN_D0 - K + 1;
    const int nt = (31 + K + 31) / 32;
    for (int j = 0; j < nt; ++j) {
        const int Xr = rb + 32 * j + nl;
        const bf16x8 b0 = load_ch8(fr, Hq, Wq, Y, Xr, 8 * g);
        const bf16x8 b1 = load_ch8(fr, Hq, Wq, Y, Xr, 16 + 8 * g);
        f32x16 acc = {0};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
        // every (cell, bin) entry comes up exactly once over the tiles:
        // a right cell past the row gives -1e30 (no cell), which the scans
        // below skip
        // (entries outside [0, K) go to a spare slot past the wave's table,
        // so the stores need no branches)
        const bool in_row = Xr >= 0 && Xr < Wq;
        const int kr = X0 - Xr - CNN_D0 + 4 * g;
        const int rbase = 4 * g * ld + kr;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int mi = 8 * (i / 4) + (i % 4);  // m - 4 g
            const int k = kr + mi;
            const int at = (unsigned)k < (unsigned)K ? rbase + mi * (ld + 1) : 32 * ld;
            S[at] = in_row ? acc[i] * 0.17677669529663687f : -1e30f;  // 1 / sqrt(32)
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // cell nl: lane half g scans bins [g * Kh, min(K, (g + 1) * Kh))
    const int Kh = (K + 1) / 2, k0 = g * Kh, k1 = min(K, k0 + Kh);
    const float* Sr = S + nl * ld;
    int kb = k0;
    float mx = -1e30f;
    for (int k = k0; k < k1; ++k) {
        const float c = Sr[k];
        if (c > mx) { mx = c; kb = k; }
    }
    // first maximum over both halves: the upper half's only if larger
    const float omx = __shfl_xor(mx, 32);
    const int okb = __shfl_xor(kb, 32);
    const float lmx = g ? omx : mx, umx = g ? mx : omx;
    const int lkb = g ? okb : kb, ukb = g ? kb : okb;
    const bool up = umx > lmx;
    mx = up ? umx : lmx;
    kb = up ? ukb : lkb;
    float z = 0.0f;
    for (int k = k0; k < k1; ++k) {
        const float c = Sr[k];
        z += c > -1e29f ? __expf(c - mx) : 0.0f;
    }
    z += __shfl_xor(z, 32);
    const int X = X0 + nl;
    if (g || X >= Wq)
        return;
    const float pk = 1.0f / z;
    float kf = (float)kb;
    if (kb > 0 && kb < K - 1) {
        const float cm = Sr[kb - 1], cp = Sr[kb + 1];
        if (cm > -1e29f && cp > -1e29f) {
            const float den = cm - 2.0f * mx + cp;
            if (den < 0.0f) kf += fminf(fmaxf(0.5f * (cm - cp) / den, -0.5f), 0.5f);
        }
    }
    const size_t at = (size_t)Y * Wq + X;
    center[at] = CNN_DS * (CNN_D0 + kf);
    halfwidth[at] = (int16_t)(pk >= p_conf ? hw_lo : hw_hi);
    if (peak_prob) peak_prob[at] = pk;
}

}  // namespace bicos_hip
