// Learned search window: CNN training and inference entry points and the
// windowed matching path (include/bicos_hip.h, "learned search window").
//
// bicos_hip_match_windowed runs BICOS::match (src/impl/cuda.cu.in:54-292
// pipeline) with the full-row search replaced by
//   CNN windows -> windowed search -> NCC stage -> fallback list ->
//   full-row search of the listed pixels -> NCC stage of the listed pixels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <string>

#include "host_common.hpp"
#include "window.hpp"

using namespace bicos_hip;

namespace bicos_hip {
namespace {

constexpr int CNN_BANDS_ROWS = 8;  // cell rows per weight-gradient band

struct Grid {
    int H, W, Hq, Wq, K;
    size_t cells;
};

int check_options(const bicos_window_options* o, int depth, float* in_scale, int* K) {
    if (!o) return fail(BICOS_ERR_INVALID_ARGUMENT, "null window options");
    if (o->bins < 9 || o->bins > 1024)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "window bins must be in [9, 1024]");
    if (o->hw_lo < 0 || o->hw_hi < 0 || o->hw_lo > 32767 || o->hw_hi > 32767)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "window half-widths must be in [0, 32767]");
    if (o->fallback != BICOS_FALLBACK_NONE && o->fallback != BICOS_FALLBACK_ALL &&
        o->fallback != BICOS_FALLBACK_UNSURE)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown fallback mode");
    if (!(o->in_scale >= 0.0f) || !std::isfinite(o->in_scale))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "bad input scale");
    *in_scale = o->in_scale > 0.0f ? o->in_scale : 8.0f / (depth == BICOS_8U ? 255.0f : 65535.0f);
    *K = o->bins;
    return BICOS_OK;
}

int check_pair(const bicos_image* stack0, const bicos_image* stack1, uint32_t n, StackInfo* i0,
               StackInfo* i1, FrameTable* f0, FrameTable* f1) {
    int rc;
    if ((rc = check_stack(stack0, n, i0, f0))) return rc;
    if ((rc = check_stack(stack1, n, i1, f1))) return rc;
    if (i0->rows != i1->rows || i0->cols != i1->cols || i0->depth != i1->depth)
        return fail(BICOS_ERR_EXCEPTION, "the two stacks differ in size or type");
    if (n < (uint32_t)CNN_IN)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "the learned window needs at least 4 frames");
    if (i0->rows < CNN_DS || i0->cols < CNN_DS)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "the learned window needs at least 4 rows and columns");
    if (i0->rows > 65535) return fail(BICOS_ERR_INVALID_ARGUMENT, "at most 65535 rows");
    return BICOS_OK;
}

/* --------------------------------------------------------------- network */

struct ViewActs {
    float *x0, *a1, *a2, *a3, *f;  // 4, 32, 32, 32, 32 planes of cells
};

size_t acts_floats(size_t cells) { return (size_t)(CNN_IN + 4 * CNN_C) * cells; }

ViewActs carve_acts(float* p, size_t cells) {
    ViewActs a;
    a.x0 = p;
    a.a1 = a.x0 + CNN_IN * cells;
    a.a2 = a.a1 + CNN_C * cells;
    a.a3 = a.a2 + CNN_C * cells;
    a.f = a.a3 + CNN_C * cells;
    return a;
}

int run_bf16_weights(const float* params, float* wbf, hipStream_t s) {
    LAUNCH("cnn_weights", k_cnn_bf16_weights, dim3((CNN_PARAMS + 255) / 256), dim3(256), 0, s, params, wbf);
    return BICOS_OK;
}

int cnn_forward(const Stack& f, int depth, const Grid& g, float in_scale, const float* wbf,
                const float* params, const ViewActs& a, hipStream_t s) {
    const dim3 grid((g.Wq + 255) / 256, g.Hq), block(256);
    if (depth == BICOS_8U)
        LAUNCH("cnn_input", k_cnn_input<uint8_t>, grid, block, 0, s, f, g.W, g.Hq, g.Wq, in_scale, a.x0);
    else
        LAUNCH("cnn_input", k_cnn_input<uint16_t>, grid, block, 0, s, f, g.W, g.Hq, g.Wq, in_scale, a.x0);
    LAUNCH("cnn_conv", (k_cnn_conv<CNN_IN, true>), grid, block, 0, s, (const float*)a.x0,
           wbf + cnn_w_off(0), params + cnn_b_off(0), g.Hq, g.Wq, a.a1);
    LAUNCH("cnn_conv", (k_cnn_conv<CNN_C, true>), grid, block, 0, s, (const float*)a.a1,
           wbf + cnn_w_off(1), params + cnn_b_off(1), g.Hq, g.Wq, a.a2);
    LAUNCH("cnn_conv", (k_cnn_conv<CNN_C, true>), grid, block, 0, s, (const float*)a.a2,
           wbf + cnn_w_off(2), params + cnn_b_off(2), g.Hq, g.Wq, a.a3);
    LAUNCH("cnn_conv", (k_cnn_conv<CNN_C, false>), grid, block, 0, s, (const float*)a.a3,
           wbf + cnn_w_off(3), params + cnn_b_off(3), g.Hq, g.Wq, a.f);
    return BICOS_OK;
}

template<int CIN>
int weight_grad(const float* x, const float* dy, const Grid& g, int layer, float* partial,
                float* grad, hipStream_t s) {
    constexpr int NT = CIN * 9 + 1;
    const int bands = (g.Hq + CNN_BANDS_ROWS - 1) / CNN_BANDS_ROWS;
    LAUNCH("cnn_weight_grad", k_cnn_conv_bwd_weight<CIN>, dim3(CNN_C, bands), dim3(256), 0, s, x, dy,
           g.Hq, g.Wq, CNN_BANDS_ROWS, partial);
    LAUNCH("cnn_weight_grad", k_cnn_reduce_weight<CIN>, dim3((CNN_C * NT + 255) / 256), dim3(256), 0, s,
           (const float*)partial, bands, layer, 1, grad);
    return BICOS_OK;
}

// backward pass of one view from d (gradient w.r.t. its features; used as
// scratch), accumulating into grad
int cnn_backward(const ViewActs& a, float* d, float* e, const float* wbf, const Grid& g,
                 float* partial, float* grad, hipStream_t s) {
    const dim3 grid((g.Wq + 255) / 256, g.Hq), block(256);
    const size_t n32 = (size_t)CNN_C * g.cells;
    const dim3 gflat((unsigned)((n32 + 255) / 256));
    int rc;
    // layer 3 (no ReLU after it)
    if ((rc = weight_grad<CNN_C>(a.a3, d, g, 3, partial, grad, s))) return rc;
    LAUNCH("cnn_data_grad", k_cnn_conv_bwd_data<CNN_C>, grid, block, 0, s, (const float*)d, wbf + cnn_w_off(3), g.Hq, g.Wq, e);
    LAUNCH("cnn_relu_grad", k_cnn_relu_bwd, gflat, block, 0, s, (const float*)a.a3, e, n32);
    // layer 2
    if ((rc = weight_grad<CNN_C>(a.a2, e, g, 2, partial, grad, s))) return rc;
    LAUNCH("cnn_data_grad", k_cnn_conv_bwd_data<CNN_C>, grid, block, 0, s, (const float*)e, wbf + cnn_w_off(2), g.Hq, g.Wq, d);
    LAUNCH("cnn_relu_grad", k_cnn_relu_bwd, gflat, block, 0, s, (const float*)a.a2, d, n32);
    // layer 1
    if ((rc = weight_grad<CNN_C>(a.a1, d, g, 1, partial, grad, s))) return rc;
    LAUNCH("cnn_data_grad", k_cnn_conv_bwd_data<CNN_C>, grid, block, 0, s, (const float*)d, wbf + cnn_w_off(1), g.Hq, g.Wq, e);
    LAUNCH("cnn_relu_grad", k_cnn_relu_bwd, gflat, block, 0, s, (const float*)a.a1, e, n32);
    // layer 0
    return weight_grad<CNN_IN>(a.x0, e, g, 0, partial, grad, s);
}

size_t partial_floats(const Grid& g) {
    return (size_t)((g.Hq + CNN_BANDS_ROWS - 1) / CNN_BANDS_ROWS) * CNN_C * (CNN_C * 9 + 1);
}

// Inference implementation of the network: matrix cores (default) or the
// training kernels on the vector ALU (for tests and measurement).
int g_cnn_impl = BICOS_CNN_MFMA;

size_t valu_scratch_bytes(const Grid& g) {
    return 4 * (align_up(CNN_PARAMS, 64) + 2 * acts_floats(g.cells) + (size_t)g.K * g.cells);
}

size_t mfma_scratch_bytes(const Grid& g) {
    return align_up(CNN_MFMA_W0 * 2, 256) + align_up(CNN_MFMA_W * 2, 256) +
           2 * align_up(CNN_IN * g.cells * 2, 256) + 4 * align_up(CNN_C * g.cells * 2, 256);
}

size_t windows_scratch_bytes(const Grid& g) {
    return std::max(valu_scratch_bytes(g), mfma_scratch_bytes(g)) + 256;
}

// Windows of all cells into center/halfwidth (peak nullable); scratch holds
// windows_scratch_bytes(g).
// zero (nullable): nzero counters the network's first kernel clears on the way
int cnn_windows(const Stack& v0, const Stack& v1, int depth, const Grid& g, float in_scale,
                const float* params, const bicos_window_options* o, void* scratch, float* center,
                int16_t* halfwidth, float* peak, hipStream_t s, int* zero = nullptr, int nzero = 0) {
    const dim3 grid((g.Wq + 255) / 256, g.Hq), block(256);
    if (g_cnn_impl == BICOS_CNN_MFMA || g_cnn_impl == BICOS_CNN_MFMA_LAYERS) {
        if (g.K > CNN_KMAX_MFMA)
            return fail(BICOS_ERR_INVALID_ARGUMENT, "window bins above 160 need the vector-ALU network");
        uint8_t* p = (uint8_t*)scratch;
        uint16_t* w0 = (uint16_t*)p; p += align_up(CNN_MFMA_W0 * 2, 256);
        uint16_t* wm = (uint16_t*)p; p += align_up(CNN_MFMA_W * 2, 256);
        uint16_t* x0L = (uint16_t*)p; p += align_up(CNN_IN * g.cells * 2, 256);
        uint16_t* x0R = (uint16_t*)p; p += align_up(CNN_IN * g.cells * 2, 256);
        uint16_t* aL = (uint16_t*)p; p += align_up(CNN_C * g.cells * 2, 256);
        uint16_t* bL = (uint16_t*)p; p += align_up(CNN_C * g.cells * 2, 256);
        uint16_t* aR = (uint16_t*)p; p += align_up(CNN_C * g.cells * 2, 256);
        uint16_t* bR = (uint16_t*)p;
        // one 4-sample load per cell row where both stacks allow it
        const size_t vb = 4 * (depth == BICOS_8U ? 1 : 2);
        const auto aligned = [&](const Stack& v) {
            return ((uintptr_t)v.base % vb) == 0 && ((size_t)v.fstride % vb) == 0 && ((size_t)v.step % vb) == 0;
        };
        const bool vec = aligned(v0) && aligned(v1);
        // both views' inputs, and the weights packed by the blocks of z = 2
        const dim3 igrid((g.Wq + 255) / 256, g.Hq, 3);
#define CI(T, V) LAUNCH("cnn_input", (k_cnn_input_pair<T, V>), igrid, block, 0, s, v0, v1, g.Hq, g.Wq, in_scale, x0L, x0R, params, w0, wm, (uint32_t*)zero, zero ? nzero : 0)
        if (depth == BICOS_8U) {
            if (vec) CI(uint8_t, true); else CI(uint8_t, false);
        } else {
            if (vec) CI(uint16_t, true); else CI(uint16_t, false);
        }
#undef CI
        const int tiles_row = (g.Wq + 31) / 32;
        if (g_cnn_impl == BICOS_CNN_MFMA) {
            const int ft = ((g.Wq + CNN_FT_W - 1) / CNN_FT_W) * ((g.Hq + CNN_FT_H - 1) / CNN_FT_H);
            LAUNCH("cnn_features", k_cnn_features_fused, dim3(ft, 1, 2), dim3(CNN_FT_THREADS), 0, s,
                   (const uint16_t*)x0L, (const uint16_t*)x0R, (const uint16_t*)w0, (const uint16_t*)wm,
                   params, g.Hq, g.Wq, bL, bR);
        } else {
        LAUNCH("cnn_conv0", k_cnn_conv0_mfma, dim3((g.Hq * tiles_row + 3) / 4, 1, 2), block, 0, s,
               (const uint16_t*)x0L, (const uint16_t*)x0R, (const uint16_t*)w0, params + cnn_b_off(0),
               g.Hq, g.Wq, aL, aR);
        const dim3 cgrid(tiles_row * ((g.Hq + CNN_TROWS - 1) / CNN_TROWS), 1, 2);
        const size_t wl = (size_t)CNN_C * 9 * CNN_C;
        LAUNCH("cnn_conv", k_cnn_conv_mfma<true>, cgrid, dim3(CNN_CONV_THREADS), 0, s, (const uint16_t*)aL, (const uint16_t*)aR,
               (const uint16_t*)wm, params + cnn_b_off(1), g.Hq, g.Wq, bL, bR);
        LAUNCH("cnn_conv", k_cnn_conv_mfma<true>, cgrid, dim3(CNN_CONV_THREADS), 0, s, (const uint16_t*)bL, (const uint16_t*)bR,
               (const uint16_t*)(wm + wl), params + cnn_b_off(2), g.Hq, g.Wq, aL, aR);
        LAUNCH("cnn_conv", k_cnn_conv_mfma<false>, cgrid, dim3(CNN_CONV_THREADS), 0, s, (const uint16_t*)aL, (const uint16_t*)aR,
               (const uint16_t*)(wm + 2 * wl), params + cnn_b_off(3), g.Hq, g.Wq, bL, bR);
        }
        const int cw = g.Hq * tiles_row;
        const size_t clds = cnn_corr_window_lds(g.K);
        if (clds > 65536)  // above the default limit of dynamic LDS per workgroup
            HIP_TRY(hipFuncSetAttribute((const void*)k_cnn_corr_window,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
        LAUNCH("cnn_corr_window", k_cnn_corr_window, dim3((cw + 3) / 4), block, clds, s, (const uint16_t*)bL,
               (const uint16_t*)bR, g.Hq, g.Wq, g.K, o->p_conf, o->hw_lo, o->hw_hi, center, halfwidth, peak);
        return BICOS_OK;
    }
    float* wbf = (float*)scratch;
    float* actp = wbf + align_up(CNN_PARAMS, 64);
    const ViewActs aL = carve_acts(actp, g.cells);
    const ViewActs aR = carve_acts(actp + acts_floats(g.cells), g.cells);
    float* C = actp + 2 * acts_floats(g.cells);
    int rc;
    if ((rc = run_bf16_weights(params, wbf, s))) return rc;
    if ((rc = cnn_forward(v0, depth, g, in_scale, wbf, params, aL, s))) return rc;
    if ((rc = cnn_forward(v1, depth, g, in_scale, wbf, params, aR, s))) return rc;
    LAUNCH("cnn_corr", k_cnn_corr, grid, block, 0, s, (const float*)aL.f, (const float*)aR.f, g.Hq, g.Wq, g.K, C);
    LAUNCH("cnn_window", k_cnn_window, grid, block, 0, s, (const float*)C, g.Hq, g.Wq, g.K, o->p_conf,
           o->hw_lo, o->hw_hi, center, halfwidth, peak);
    return BICOS_OK;
}

/* --------------------------------------------------------- side stream */

// Per device: a second stream (default priority) on which the descriptor
// transform runs while the network runs on the caller's stream, and the
// fork / join events. Guarded by g_ws_mu (held for the whole call).
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
std::map<int, SideStream> g_side;

int side_stream(int dev, SideStream** out) {
    SideStream& ss = g_side[dev];
    if (!ss.stream) {
        HIP_TRY(hipStreamCreateWithFlags(&ss.stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
    }
    *out = &ss;
    return BICOS_OK;
}

/* ------------------------------------------------------- windowed search */

template<int WORDS>
int run_window_search_w(const uint32_t* d0, const uint32_t* d1, const Grid& g, bool nodupes,
                        bool consist, int max_lr, const float* center, const int16_t* hw,
                        int16_t* out, int64_t ostep, hipStream_t s) {
    const dim3 grid((g.W + 255) / 256, g.H), block(256);
#define WS(ND, CS) LAUNCH("window_search", (k_window_search<WORDS, ND, CS>), grid, block, 0, s, d0, d1, g.W, center, hw, g.Hq, g.Wq, max_lr, out, ostep)
    if (nodupes) {
        if (consist) WS(true, true); else WS(true, false);
    } else {
        if (consist) WS(false, true); else WS(false, false);
    }
#undef WS
    return BICOS_OK;
}

int run_window_search(const uint32_t* d0, const uint32_t* d1, const Grid& g, int words,
                      bool nodupes, bool consist, int max_lr, const float* center,
                      const int16_t* hw, int16_t* out, int64_t ostep, hipStream_t s) {
    switch (words) {
        case 1: return run_window_search_w<1>(d0, d1, g, nodupes, consist, max_lr, center, hw, out, ostep, s);
        case 2: return run_window_search_w<2>(d0, d1, g, nodupes, consist, max_lr, center, hw, out, ostep, s);
        case 4: return run_window_search_w<4>(d0, d1, g, nodupes, consist, max_lr, center, hw, out, ostep, s);
        case 8: return run_window_search_w<8>(d0, d1, g, nodupes, consist, max_lr, center, hw, out, ostep, s);
        default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
    }
}

template<int WORDS>
int run_search_list_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes, int mode,
                      const int16_t* qcol, const int* row_count, int16_t* res, int16_t* disp,
                      int64_t dstep, hipStream_t s) {
    const int chunk = (int)std::min<size_t>((size_t)W, 64 * 1024 / (WORDS * 4));
    const size_t lds = (size_t)chunk * WORDS * 4;
    const dim3 grid(H, (W + SEARCH_THREADS - 1) / SEARCH_THREADS), block(SEARCH_THREADS);
#define SL(ND, M) LAUNCH("search_list", (k_search_list<WORDS, ND, M>), grid, block, lds, s, dq, dc, W, chunk, qcol, row_count, res, disp, dstep)
    if (nodupes) {
        if (mode == SEARCH_REV) SL(true, SEARCH_REV); else SL(true, SEARCH_DISP);
    } else {
        if (mode == SEARCH_REV) SL(false, SEARCH_REV); else SL(false, SEARCH_DISP);
    }
#undef SL
    return BICOS_OK;
}

template<int WORDS, int NB, int WAVES, int CH>
int run_search_list_mfma_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes,
                           int mode, const int16_t* qcol, const int* row_count, int16_t* res,
                           int16_t* disp, int64_t dstep, hipStream_t s) {
    constexpr size_t lds = (size_t)CH * (2 * WORDS + 1) * 16;
    const dim3 grid(H, (W + WAVES * NB * 32 - 1) / (WAVES * NB * 32)), block(WAVES * 64);
#define SM(ND, M, E) LAUNCH("search_list_mfma", (k_search_mfma<WORDS, NB, ND, M, WAVES, CH, true, E>), grid, block, lds, s, dq, dc, W, (const int16_t*)nullptr, 0, disp, dstep, qcol, row_count, res)
#define SM2(ND, M) do { if (W <= EMBED_MAX_W) SM(ND, M, true); else SM(ND, M, false); } while (0)
    if (nodupes) {
        if (mode == SEARCH_REV) SM2(true, SEARCH_REV); else SM2(true, SEARCH_DISP);
    } else {
        if (mode == SEARCH_REV) SM2(false, SEARCH_REV); else SM2(false, SEARCH_DISP);
    }
#undef SM2
#undef SM
    return BICOS_OK;
}

template<int WORDS, int NB, int WAVES, int CH>
int run_search_list_mfma4_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes,
                            int mode, const int16_t* qcol, const int* row_count, int16_t* res,
                            int16_t* disp, int64_t dstep, hipStream_t s) {
    constexpr size_t lds = (size_t)CH * (WORDS + 1) * 16;
    // two blocks per row, taking alternate groups of entries (a row lists a
    // tenth of its pixels or so: mostly a single group, which the second
    // block leaves at once; rows listing more do not wait on one block)
    const dim3 grid(H, 2), block(WAVES * 64);
#define SM(ND, M) do { \
        int rc_ = allow_dynamic_lds((const void*)k_search_mfma_fp4<WORDS, NB, ND, M, WAVES, CH, true>, lds); \
        if (rc_) return rc_; \
        LAUNCH("search_list_mfma", (k_search_mfma_fp4<WORDS, NB, ND, M, WAVES, CH, true>), grid, block, lds, s, dq, dc, W, (const int16_t*)nullptr, 0, disp, dstep, qcol, row_count, res); \
    } while (0)
    if (nodupes) {
        if (mode == SEARCH_REV) SM(true, SEARCH_REV); else SM(true, SEARCH_DISP);
    } else {
        if (mode == SEARCH_REV) SM(false, SEARCH_REV); else SM(false, SEARCH_DISP);
    }
#undef SM
    return BICOS_OK;
}

// Full-row search of the listed pixels: the full search's kernels with one
// 32-query tile per wave (a row lists a tenth of its pixels or so, and
// smaller query groups keep more of a workgroup's waves busy). 128-bit rows
// are staged in chunks of 480 columns (38 KB of LDS, three workgroups per
// CU): with so few queries per row, the staging's latency is what a
// workgroup waits on, and a third workgroup per CU hides more of it.
// 256-bit rows take 512 columns (72 KiB, two workgroups per CU; a
// 4096-column row is eight chunks): 0.870 ms at 4096x3000x64 against 0.88
// for 448 columns and 0.90 for 352 (three workgroups per CU).
int run_search_list(const uint32_t* dq, const uint32_t* dc, int H, int W, int words, bool nodupes,
                    int mode, const int16_t* qcol, const int* row_count, int16_t* res,
                    int16_t* disp, int64_t dstep, hipStream_t s) {
    if (g_search_impl == BICOS_SEARCH_MFMA && words >= 2 && W <= EMBED_MAX_W) {
        switch (words) {
            case 2: return run_search_list_mfma4_w<2, 1, 8, 1344>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            case 4: return run_search_list_mfma4_w<4, 1, 8, 480>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            case 8: return run_search_list_mfma4_w<8, 1, 8, 512>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
        }
    }
    if (g_search_impl == BICOS_SEARCH_MFMA || g_search_impl == BICOS_SEARCH_MFMA_I8) {
        switch (words) {
            case 1: return run_search_list_mfma_w<1, 1, 8, 1344>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            case 2: return run_search_list_mfma_w<2, 1, 8, 800>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            case 4: return run_search_list_mfma_w<4, 1, 8, 448>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            case 8: return run_search_list_mfma_w<8, 1, 8, 224>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
            default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
        }
    }
    switch (words) {
        case 1: return run_search_list_w<1>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
        case 2: return run_search_list_w<2>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
        case 4: return run_search_list_w<4>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
        case 8: return run_search_list_w<8>(dq, dc, H, W, nodupes, mode, qcol, row_count, res, disp, dstep, s);
        default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
    }
}

__global__ void k_zero_u32(unsigned int* p, int count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// status non-null (MOMENTS): the set from the NCC stage's status bits;
// hw non-null: only pixels of cells given the wide window
int run_fallback_list(const float* disp_f, const int16_t* disp_i, int64_t dstep, const Stack& v0,
                      int depth, uint32_t n, const Grid& g, int has_mv, float mv_n, float thr,
                      int precision, const uint8_t* status, const int16_t* hw, int hw_lo,
                      int16_t* list, int* row_count, hipStream_t s) {
    LAUNCH("fallback_list", k_zero_u32, dim3((g.H + 255) / 256), dim3(256), 0, s, (unsigned int*)row_count, g.H);
    const dim3 grid(g.H, (g.W + 255) / 256), block(256);
    if (status) {
        LAUNCH("fallback_list", k_fallback_list_status, grid, block, 0, s, status, disp_f, dstep, g.W,
               hw, hw_lo, g.Hq, g.Wq, list, row_count);
        return run_row_scan(row_count, g.H, s);
    }
#define FL(TI, TP) LAUNCH("fallback_list", (k_fallback_list<TI, TP>), grid, block, 0, s, disp_f, disp_i, dstep, v0, (int)n, g.W, has_mv, mv_n, thr, hw, hw_lo, g.Hq, g.Wq, list, row_count)
    if (precision == BICOS_PRECISION_DOUBLE) {
        if (depth == BICOS_8U) FL(uint8_t, double); else FL(uint16_t, double);
    } else {
        if (depth == BICOS_8U) FL(uint8_t, float); else FL(uint16_t, float);
    }
#undef FL
    return run_row_scan(row_count, g.H, s);
}

// The fused windowed pass of the MOMENTS_FAST pipeline (k_window_core)
template<typename TIn, int WORDS>
int run_window_core_w(const uint32_t* d0, const uint32_t* d1, const Grid& g, bool nodupes,
                      bool consist, int max_lr, const float* center, const int16_t* hw,
                      const Stack& v0, const Stack& v1, uint32_t n, float thr, float step,
                      int has_mv, float mv_n, float* out, int64_t ostep, int fallback, int hw_lo,
                      int16_t* list, int* row_count, int16_t* dwin, hipStream_t s) {
    const GridX gx = grid_x(step);
    const int npts = subpixel_points(step);
    const dim3 grid((g.W + 255) / 256, g.H), block(256);
    const int fb = fallback != BICOS_FALLBACK_NONE, unsure = fallback == BICOS_FALLBACK_UNSURE;
    // staged span: at most 1024 columns and 32 KiB of LDS per block
    const int span = std::min(1024, (int)((32u << 10) / core_lds_bytes(WORDS, consist, 1))) & ~7;
    const size_t lds = core_lds_bytes(WORDS, consist, span);
#define WC(ND, CS) do { \
        int rc_ = allow_dynamic_lds((const void*)k_window_core<TIn, WORDS, ND, CS>, lds); \
        if (rc_) return rc_; \
        LAUNCH("window_core", (k_window_core<TIn, WORDS, ND, CS>), grid, block, lds, s, d0, d1, center, hw, g.Hq, g.Wq, max_lr, v0, v1, (int)n, g.W, thr, npts, gx, has_mv, mv_n, out, ostep, fb, unsure, hw_lo, list, row_count, span, dwin); \
    } while (0)
    if (!consist) WC(true, false);
    else if (nodupes) WC(true, true);
    else WC(false, true);
#undef WC
    return BICOS_OK;
}

template<typename TIn>
int run_window_core_t(const uint32_t* d0, const uint32_t* d1, const Grid& g, int words, bool nodupes,
                      bool consist, int max_lr, const float* center, const int16_t* hw,
                      const Stack& v0, const Stack& v1, uint32_t n, float thr, float step,
                      int has_mv, float mv_n, float* out, int64_t ostep, int fallback, int hw_lo,
                      int16_t* list, int* row_count, int16_t* dwin, hipStream_t s) {
#define WCW(WD) return run_window_core_w<TIn, WD>(d0, d1, g, nodupes, consist, max_lr, center, hw, v0, v1, n, thr, step, has_mv, mv_n, out, ostep, fallback, hw_lo, list, row_count, dwin, s)
    switch (words) {
        case 1: WCW(1);
        case 2: WCW(2);
        case 4: WCW(4);
        case 8: WCW(8);
        default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
    }
#undef WCW
}

}  // namespace
}  // namespace bicos_hip

/* ================================================================== C ABI */

extern "C" {

int bicos_hip_window_param_count(void) { return CNN_PARAMS; }

int bicos_hip_set_cnn_impl(int32_t impl) {
    if (impl != BICOS_CNN_MFMA && impl != BICOS_CNN_VALU && impl != BICOS_CNN_MFMA_LAYERS)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown network implementation");
    g_cnn_impl = impl;
    return BICOS_OK;
}

void bicos_hip_window_options_default(bicos_window_options* o) {
    if (!o) return;
    o->in_scale = 0.0f;
    o->p_conf = 0.5f;
    o->hw_lo = 6;
    o->hw_hi = 12;
    o->bins = CNN_K;
    o->fallback = BICOS_FALLBACK_ALL;
}

int bicos_hip_window_init(float* p, uint64_t seed) {
    if (!p) return fail(BICOS_ERR_INVALID_ARGUMENT, "null parameter buffer");
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 4; ++l) {
        const int cin = l == 0 ? CNN_IN : CNN_C;
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (cin * 9)));
        for (int i = 0; i < CNN_C * cin * 9; ++i) p[cnn_w_off(l) + i] = (float)nd(rng);
        for (int i = 0; i < CNN_C; ++i) p[cnn_b_off(l) + i] = 0.0f;
    }
    return BICOS_OK;
}

int bicos_hip_window_label(const bicos_image* disparity, int32_t min_valid, float* label,
                           void* stream) {
    if (!disparity || !disparity->data || !label) return fail(BICOS_ERR_INVALID_ARGUMENT, "null buffer");
    if (disparity->type != BICOS_32F && disparity->type != BICOS_16S)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "disparity must be BICOS_32F or BICOS_16S");
    if (disparity->rows < CNN_DS || disparity->cols < CNN_DS)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "disparity smaller than one cell");
    const int es = elem_size(disparity->type);
    if (disparity->step % es || disparity->step < (int64_t)disparity->cols * es)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "bad row step");
    const int Hq = disparity->rows / CNN_DS, Wq = disparity->cols / CNN_DS;
    const bool f = disparity->type == BICOS_32F;
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("window_label", k_window_label, dim3((Wq + 255) / 256, Hq), dim3(256), 0, s,
           f ? (const float*)disparity->data : nullptr, f ? nullptr : (const int16_t*)disparity->data,
           disparity->step / es, Hq, Wq, (int)min_valid, label);
    return BICOS_OK;
}

int bicos_hip_window_loss_grad(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                               const float* params, const bicos_window_options* opts,
                               const float* label, float* grad, float* loss, void* stream) {
    StackInfo i0, i1;
    FrameTable f0, f1;
    int rc;
    if ((rc = check_pair(stack0, stack1, n, &i0, &i1, &f0, &f1))) return rc;
    if (!params || !label || !grad || !loss) return fail(BICOS_ERR_INVALID_ARGUMENT, "null buffer");
    float in_scale;
    Grid g;
    if ((rc = check_options(opts, i0.depth, &in_scale, &g.K))) return rc;
    g.H = i0.rows; g.W = i0.cols; g.Hq = g.H / CNN_DS; g.Wq = g.W / CNN_DS;
    g.cells = (size_t)g.Hq * g.Wq;
    hipStream_t s = (hipStream_t)stream;

    const size_t pk0 = align_up(packed_bytes(f0, n, i0), 256), pk1 = align_up(packed_bytes(f1, n, i1), 256);
    const size_t floats = align_up(CNN_PARAMS, 64) + 2 * acts_floats(g.cells) +
                          2 * (size_t)g.K * g.cells + 2 * g.cells + 4 * (size_t)CNN_C * g.cells +
                          partial_floats(g) + 64;
    int dev;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws;
    if ((rc = ws_acquire(dev, floats * 4 + pk0 + pk1 + 1024, s, &ws, WS_WINDOW))) return rc;
    auto body = [&]() -> int {
        int r;
        uint8_t* pb = (uint8_t*)ws;
        Stack v0, v1;
        if ((r = make_stack(f0, n, i0, pb, &v0, s))) return r;
        if ((r = make_stack(f1, n, i1, pb + pk0, &v1, s))) return r;
        float* p = (float*)(pb + pk0 + pk1);
        float* wbf = p; p += align_up(CNN_PARAMS, 64);
        const ViewActs aL = carve_acts(p, g.cells); p += acts_floats(g.cells);
        const ViewActs aR = carve_acts(p, g.cells); p += acts_floats(g.cells);
        float* C = p; p += (size_t)g.K * g.cells;
        float* dC = p; p += (size_t)g.K * g.cells;
        float* cell_loss = p; p += 2 * g.cells;
        float* dL = p; p += (size_t)CNN_C * g.cells;
        float* dR = p; p += (size_t)CNN_C * g.cells;
        float* e0 = p; p += (size_t)CNN_C * g.cells;
        float* e1 = p; p += (size_t)CNN_C * g.cells;
        float* partial = p;
        if ((r = run_bf16_weights(params, wbf, s))) return r;
        if ((r = cnn_forward(v0, i0.depth, g, in_scale, wbf, params, aL, s))) return r;
        if ((r = cnn_forward(v1, i0.depth, g, in_scale, wbf, params, aR, s))) return r;
        const dim3 grid((g.Wq + 255) / 256, g.Hq), block(256);
        LAUNCH("cnn_corr", k_cnn_corr, grid, block, 0, s, (const float*)aL.f, (const float*)aR.f, g.Hq, g.Wq, g.K, C);
        LAUNCH("cnn_loss", k_cnn_softmax_ce, grid, block, 0, s, (const float*)C, label, g.Hq, g.Wq, g.K, dC, cell_loss);
        LAUNCH("cnn_loss", k_loss_sum, dim3(1), block, 0, s, (const float*)cell_loss, g.cells, loss);
        LAUNCH("cnn_corr_grad", k_cnn_corr_bwd, grid, block, 0, s, (const float*)aL.f, (const float*)aR.f,
               (const float*)dC, 1.0f, g.Hq, g.Wq, g.K, dL, dR);
        if ((r = cnn_backward(aL, dL, e0, wbf, g, partial, grad, s))) return r;
        return cnn_backward(aR, dR, e1, wbf, g, partial, grad, s);
    };
    rc = body();
    const int rc2 = ws_release(dev, s, WS_WINDOW);
    return rc ? rc : rc2;
}

int bicos_hip_window_adam(float* params, float* m, float* v, const float* grad, const float* count,
                          float lr, float beta1, float beta2, float eps, int32_t step,
                          void* stream) {
    if (!params || !m || !v || !grad) return fail(BICOS_ERR_INVALID_ARGUMENT, "null buffer");
    if (step < 1) return fail(BICOS_ERR_INVALID_ARGUMENT, "Adam steps count from 1");
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("cnn_adam", k_cnn_adam, dim3((CNN_PARAMS + 255) / 256), dim3(256), 0, s, params, m, v, grad,
           count, lr, beta1, beta2, eps, c1, c2);
    return BICOS_OK;
}

int bicos_hip_window_predict(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                             const float* params, const bicos_window_options* opts, float* center,
                             int16_t* halfwidth, float* peak, void* stream) {
    StackInfo i0, i1;
    FrameTable f0, f1;
    int rc;
    if ((rc = check_pair(stack0, stack1, n, &i0, &i1, &f0, &f1))) return rc;
    if (!params || !center || !halfwidth) return fail(BICOS_ERR_INVALID_ARGUMENT, "null buffer");
    float in_scale;
    Grid g;
    if ((rc = check_options(opts, i0.depth, &in_scale, &g.K))) return rc;
    g.H = i0.rows; g.W = i0.cols; g.Hq = g.H / CNN_DS; g.Wq = g.W / CNN_DS;
    g.cells = (size_t)g.Hq * g.Wq;
    hipStream_t s = (hipStream_t)stream;
    const size_t pk0 = align_up(packed_bytes(f0, n, i0), 256), pk1 = align_up(packed_bytes(f1, n, i1), 256);
    int dev;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws;
    if ((rc = ws_acquire(dev, windows_scratch_bytes(g) + pk0 + pk1 + 1024, s, &ws, WS_WINDOW))) return rc;
    auto body = [&]() -> int {
        int r;
        uint8_t* pb = (uint8_t*)ws;
        Stack v0, v1;
        if ((r = make_stack(f0, n, i0, pb, &v0, s))) return r;
        if ((r = make_stack(f1, n, i1, pb + pk0, &v1, s))) return r;
        return cnn_windows(v0, v1, i0.depth, g, in_scale, params, opts, pb + pk0 + pk1, center,
                           halfwidth, peak, s);
    };
    rc = body();
    const int rc2 = ws_release(dev, s, WS_WINDOW);
    return rc ? rc : rc2;
}

int bicos_hip_search_windowed(const uint32_t* desc0, const uint32_t* desc1, int32_t rows,
                              int32_t cols, int32_t words, int32_t flags, int32_t max_lr_diff,
                              const float* center, const int16_t* halfwidth, int16_t* out,
                              void* stream) {
    if (!desc0 || !desc1 || !center || !halfwidth || !out)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "null buffer");
    if (rows < CNN_DS || cols < CNN_DS)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "the learned window needs at least 4 rows and columns");
    if (rows > 65535 || cols > 32767) return fail(BICOS_ERR_INVALID_ARGUMENT, "image too large");
    if (flags < 0 || flags > 3) return fail(BICOS_ERR_INVALID_ARGUMENT, "flags must be 0..3");
    Grid g;
    g.H = rows; g.W = cols; g.Hq = rows / CNN_DS; g.Wq = cols / CNN_DS; g.K = 0;
    g.cells = (size_t)g.Hq * g.Wq;
    return run_window_search(desc0, desc1, g, words, (flags & 1) != 0, (flags & 2) != 0,
                             max_lr_diff, center, halfwidth, out, cols, (hipStream_t)stream);
}

int bicos_hip_match_windowed(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                             const bicos_config* cfg, const float* params,
                             const bicos_window_options* opts, bicos_image* disparity,
                             bicos_image* corrmap, bicos_window_stats* stats, void* stream) {
    if (!cfg) return fail(BICOS_ERR_INVALID_ARGUMENT, "null config");
    if (!params) return fail(BICOS_ERR_INVALID_ARGUMENT, "null window parameters");
    StackInfo i0, i1;
    FrameTable f0, f1;
    int rc;
    if ((rc = check_pair(stack0, stack1, n, &i0, &i1, &f0, &f1))) return rc;
    if (cfg->mode != BICOS_MODE_LIMITED && cfg->mode != BICOS_MODE_FULL)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown transform mode");
    if (cfg->variant != BICOS_VARIANT_NO_DUPLICATES && cfg->variant != BICOS_VARIANT_CONSISTENCY)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented");
    if (!precision_known(cfg->precision))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented precision");
    float in_scale;
    Grid g;
    if ((rc = check_options(opts, i0.depth, &in_scale, &g.K))) return rc;
    const int bits = required_bits(n, cfg->mode);
    const int words = words_for_bits(bits);
    if (!words)
        return fail(BICOS_ERR_INVALID_ARGUMENT,
                    "input stacks too large, would require " + std::to_string(bits) + " bits");
    const bool thr = cfg->has_nxcorr_threshold != 0;
    const bool subpix = thr && cfg->has_subpixel_step;
    if (subpix && (rc = check_step(cfg->subpixel_step))) return rc;
    g.H = i0.rows; g.W = i0.cols; g.Hq = g.H / CNN_DS; g.Wq = g.W / CNN_DS;
    g.cells = (size_t)g.Hq * g.Wq;
    const int H = g.H, W = g.W;
    if ((rc = check_out(disparity, H, W, bicos_hip_disparity_type(cfg), "disparity"))) return rc;
    if (corrmap && thr && (rc = check_out(corrmap, H, W, bicos_hip_corrmap_type(cfg), "corrmap")))
        return rc;
    hipStream_t s = (hipStream_t)stream;

    const bool consistency = cfg->variant == BICOS_VARIANT_CONSISTENCY;
    const bool nodupes = consistency ? cfg->no_dupes != 0 : true;  // cpu.cpp:70-77
    const int max_lr = consistency ? cfg->max_lr_diff : -1;
    const bool fallback = opts->fallback != 0 && thr;
    const size_t npx = (size_t)W * H;
    const size_t desc_bytes = align_up(npx * words * 4, 256);
    const size_t raw_bytes = subpix ? align_up(npx * 2, 256) : 0;
    const size_t cell_bytes = align_up(g.cells * 4, 256) + align_up(g.cells * 2, 256);
    const bool moments = subpix && cfg->precision == BICOS_PRECISION_MOMENTS;
    const size_t fb_bytes = fallback ? align_up(npx * 2, 256) + list_counts_bytes(H) + 256 +
                                           (consistency ? 2 * align_up(npx * 2, 256) : 0)
                                     : 0;
    const size_t mom_bytes = moments ? MomentsScratch::bytes(npx) : 0;
    const size_t cnn_bytes = align_up(windows_scratch_bytes(g), 256);
    const size_t pk0 = align_up(packed_bytes(f0, n, i0), 256), pk1 = align_up(packed_bytes(f1, n, i1), 256);

    int dev;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws;
    if ((rc = ws_acquire(dev, 2 * desc_bytes + raw_bytes + cell_bytes + fb_bytes + mom_bytes + cnn_bytes + pk0 + pk1, s, &ws)))
        return rc;
    uint8_t* p = (uint8_t*)ws;
    uint32_t* desc0 = (uint32_t*)p; p += desc_bytes;
    uint32_t* desc1 = (uint32_t*)p; p += desc_bytes;
    int16_t* raw = subpix ? (int16_t*)p : nullptr; p += raw_bytes;
    float* center = (float*)p; p += align_up(g.cells * 4, 256);
    int16_t* hw = (int16_t*)p; p += align_up(g.cells * 2, 256);
    int16_t *list = nullptr, *fwd = nullptr, *rev = nullptr;
    int* row_count = nullptr;
    unsigned long long* fb_total = nullptr;
    if (fallback) {
        list = (int16_t*)p; p += align_up(npx * 2, 256);
        row_count = (int*)p; p += list_counts_bytes(H);
        fb_total = (unsigned long long*)p; p += 256;
        if (consistency) {
            fwd = (int16_t*)p; p += align_up(npx * 2, 256);
            rev = (int16_t*)p; p += align_up(npx * 2, 256);
        }
    }
    const MomentsScratch ms = MomentsScratch::carve(p);
    p += mom_bytes;
    void* cnn_scratch = p; p += cnn_bytes;
    uint8_t* pack0 = p;
    uint8_t* pack1 = p + pk0;

    // MOMENTS_FAST without a corrmap: the windowed search, the subpixel stage
    // and the fallback set in one pass (k_window_core), then the fallback's
    // full-row search and its subpixel stage on the list
    const int npts_fast = subpix ? subpixel_points(cfg->subpixel_step) : 0;
    const bool fast = subpix && cfg->precision == BICOS_PRECISION_MOMENTS_FAST && !corrmap &&
                      npts_fast >= 1 && npts_fast <= 64;
    auto body = [&]() -> int {
        int r;
        Stack v0, v1;
        if ((r = make_stack(f0, n, i0, pack0, &v0, s))) return r;
        if ((r = make_stack(f1, n, i1, pack1, &v1, s))) return r;
        // the network (reads the frames only) on the caller's stream, the
        // descriptor transform beside it on the side stream. The network is
        // the longer of the two: it is submitted first, and the caller's
        // stream, which goes on with the windowed pass, finds the shorter
        // transform done when it reaches the join.
        SideStream* ss;
        if ((r = side_stream(dev, &ss))) return r;
        HIP_TRY(hipEventRecord(ss->fork, s));
        HIP_TRY(hipStreamWaitEvent(ss->stream, ss->fork, 0));
        // the fused pass appends to the row lists, which the network's first
        // kernel clears (the vector-ALU network: one more kernel)
        const bool clear = fast && fallback;
        int rc_zero = BICOS_OK;
        if (clear && g_cnn_impl == BICOS_CNN_VALU)
            rc_zero = launch("fallback_list", k_zero_u32, dim3((H + 255) / 256), dim3(256), 0, s,
                             (unsigned int*)row_count, H);
        const int rc_cnn = rc_zero ? rc_zero
                                   : cnn_windows(v0, v1, i0.depth, g, in_scale, params, opts, cnn_scratch,
                                                 center, hw, nullptr, s,
                                                 clear && g_cnn_impl != BICOS_CNN_VALU ? row_count : nullptr, H);
        r = rc_cnn ? rc_cnn
                   : transform_any(v0, &v1, i0.depth, n, cfg->mode, W, H, words, desc0, desc1, ss->stream);
        // joined in every case: the workspace is released on the caller's stream
        HIP_TRY(hipEventRecord(ss->join, ss->stream));
        HIP_TRY(hipStreamWaitEvent(s, ss->join, 0));
        if (r) return r;
        int16_t* d16 = subpix ? raw : (int16_t*)disparity->data;
        const int64_t d16_step = subpix ? W : disparity->step / 2;
        if (fast) {
            const float mv_n = cfg->has_min_variance ? cfg->min_variance * (float)n : 0.0f;
            float* outf = (float*)disparity->data;
            const int64_t ostep = disparity->step / 4;
            const int fbm = fallback ? opts->fallback : BICOS_FALLBACK_NONE;
            r = i0.depth == BICOS_8U
                ? run_window_core_t<uint8_t>(desc0, desc1, g, words, nodupes, consistency, max_lr, center, hw, v0, v1, n,
                                             cfg->nxcorr_threshold, cfg->subpixel_step, cfg->has_min_variance, mv_n,
                                             outf, ostep, fbm, opts->hw_lo, list, row_count, d16, s)
                : run_window_core_t<uint16_t>(desc0, desc1, g, words, nodupes, consistency, max_lr, center, hw, v0, v1, n,
                                              cfg->nxcorr_threshold, cfg->subpixel_step, cfg->has_min_variance, mv_n,
                                              outf, ostep, fbm, opts->hw_lo, list, row_count, d16, s);
            if (r || !fallback) return r;
            if (consistency) {
                if ((r = run_search_list(desc0, desc1, H, W, words, nodupes, SEARCH_REV, list, row_count,
                                         fwd, nullptr, 0, s)))
                    return r;
                if ((r = run_search_list(desc1, desc0, H, W, words, nodupes, SEARCH_REV, fwd, row_count,
                                         rev, nullptr, 0, s)))
                    return r;
                LAUNCH("lr_combine", k_lr_combine, dim3(H, 1), dim3(256), 0, s,
                       (const int16_t*)list, (const int*)row_count, W, (const int16_t*)fwd,
                       (const int16_t*)rev, max_lr, d16, d16_step);
            } else if ((r = run_search_list(desc0, desc1, H, W, words, nodupes, SEARCH_DISP, list, row_count,
                                            nullptr, d16, d16_step, s))) {
                return r;
            }
            if (stats)
                LAUNCH("fallback_list", k_sum_counts, dim3(1), dim3(256), 0, s, (const int*)row_count, H, fb_total);
            return run_subpixel_fast_list(d16, d16_step, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                                          cfg->subpixel_step, cfg->has_min_variance, mv_n, outf, ostep, list,
                                          row_count, s);
        }
        if ((r = run_window_search(desc0, desc1, g, words, nodupes, consistency, max_lr, center, hw,
                                   d16, d16_step, s)))
            return r;
        if (!thr) return BICOS_OK;

        const float mv_n = cfg->has_min_variance ? cfg->min_variance * (float)n : 0.0f;
        void* cm = corrmap ? corrmap->data : nullptr;
        const int64_t cstep = cm ? corrmap->step / elem_size(corrmap->type) : 0;
        float* outf = subpix ? (float*)disparity->data : nullptr;
        const int64_t ostep = subpix ? disparity->step / 4 : 0;
        if (subpix)
            r = run_subpixel(raw, W, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                             cfg->subpixel_step, cfg->has_min_variance, mv_n, cfg->precision, outf,
                             ostep, cm, cstep, s, moments ? &ms : nullptr);
        else
            r = run_agree(d16, d16_step, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                          cfg->has_min_variance, mv_n, cfg->precision, cm, cstep, s);
        if (r || !fallback) return r;

        // fallback: full-row search and NCC stage of the rejected pixels
        const bool unsure = opts->fallback == BICOS_FALLBACK_UNSURE;
        if ((r = run_fallback_list(outf, subpix ? nullptr : d16, subpix ? ostep : d16_step, v0,
                                   i0.depth, n, g, cfg->has_min_variance, mv_n,
                                   cfg->nxcorr_threshold, cfg->precision,
                                   moments && !cm ? ms.status : nullptr, unsure ? hw : nullptr,
                                   opts->hw_lo, list, row_count, s)))
            return r;
        if (consistency) {
            if ((r = run_search_list(desc0, desc1, H, W, words, nodupes, SEARCH_REV, list, row_count,
                                     fwd, nullptr, 0, s)))
                return r;
            if ((r = run_search_list(desc1, desc0, H, W, words, nodupes, SEARCH_REV, fwd, row_count,
                                     rev, nullptr, 0, s)))
                return r;
            LAUNCH("lr_combine", k_lr_combine, dim3(H, 1), dim3(256), 0, s,
                   (const int16_t*)list, (const int*)row_count, W, (const int16_t*)fwd,
                   (const int16_t*)rev, max_lr, d16, d16_step);
        } else {
            if ((r = run_search_list(desc0, desc1, H, W, words, nodupes, SEARCH_DISP, list, row_count,
                                     nullptr, d16, d16_step, s)))
                return r;
        }
        if (stats)
            LAUNCH("fallback_list", k_sum_counts, dim3(1), dim3(256), 0, s, (const int*)row_count, H, fb_total);
        if (moments && !cm)
            return run_moments_listed(raw, W, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                                      cfg->subpixel_step, cfg->has_min_variance, mv_n, outf, ostep,
                                      &ms, list, row_count, s);
        if (subpix)
            return run_subpixel_list(raw, W, v0, v1, i0.depth, n, W, cfg->nxcorr_threshold,
                                     cfg->subpixel_step, cfg->has_min_variance, mv_n,
                                     cfg->precision, outf, ostep, cm, cstep, list, row_count, H, s);
        return run_agree_list(d16, d16_step, v0, v1, i0.depth, n, W, cfg->nxcorr_threshold,
                              cfg->has_min_variance, mv_n, cfg->precision, cm, cstep, list,
                              row_count, H, s);
    };
    rc = body();
    unsigned long long fb_count = 0;
    if (!rc && stats && fallback) {
        const hipError_t e1 = hipMemcpyAsync(&fb_count, fb_total, 8, hipMemcpyDeviceToHost, s);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(s) : e1;
        if (e2 != hipSuccess) rc = fail(BICOS_ERR_HIP, std::string("fallback count: ") + hipGetErrorString(e2));
    }
    const int rc2 = ws_release(dev, s);
    if (!rc && stats) {
        stats->pixels = (int64_t)npx;
        stats->fallback_pixels = (int64_t)fb_count;
    }
    return rc ? rc : rc2;
}

}  // extern "C"
