// BICOS matching on MI355X (gfx950): host driver and C ABI.
//
// Pipeline of BICOS::match (reference: src/impl/cuda.cu.in:54-292,
// src/impl/cpu.cpp:37-155):
//   1. descriptor transform of both views        -> two descriptor planes
//   2. Hamming row search (+ reverse pass for the Consistency variant)
//   3. NCC agreement, or subpixel agreement when a step is set
// All launches go to the caller's stream. Scratch memory (descriptor planes,
// the int16 raw disparity of the subpixel path, the reverse-search columns)
// lives in one per-device workspace kept between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bicos_hip.h"
#include "agree.hpp"
#include "agree_half.hpp"
#include "agree_fast.hpp"
#include "agree_moments.hpp"
#include "search.hpp"
#include "search_mfma.hpp"
#include "host_common.hpp"
#include "transform.hpp"

using namespace bicos_hip;

namespace bicos_hip {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

/* ----------------------------------------------------------------- timing */

std::mutex g_timing_mu;
bool g_timing = false;
std::vector<TimedLaunch> g_timed;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

/* -------------------------------------------------------------- workspace */

struct Workspace {
    void* buf = nullptr;
    size_t cap = 0;
    hipEvent_t last = nullptr;
    hipStream_t last_stream = nullptr;  // stream of the last use
    bool used = false;
};

std::mutex g_ws_mu;
std::map<int, Workspace> g_ws;

int ws_acquire(int dev, size_t bytes, hipStream_t s, void** out, int slot) {
    Workspace& w = g_ws[dev * WS_SLOTS + slot];
    if (!w.last) HIP_TRY(hipEventCreateWithFlags(&w.last, hipEventDisableTiming));
    if (w.cap < bytes) {
        if (w.used) HIP_TRY(hipEventSynchronize(w.last));
        if (w.buf) HIP_TRY(hipFree(w.buf));
        w.buf = nullptr;
        w.cap = 0;
        size_t want = bytes + bytes / 8;
        HIP_TRY(hipMalloc(&w.buf, want));
        w.cap = want;
        w.used = false;
    }
    // the last use on another stream must finish first (on the same stream,
    // stream order already guarantees it)
    if (w.used && w.last_stream != s) HIP_TRY(hipStreamWaitEvent(s, w.last, 0));
    *out = w.buf;
    return BICOS_OK;
}

int ws_release(int dev, hipStream_t s, int slot) {
    Workspace& w = g_ws[dev * WS_SLOTS + slot];
    HIP_TRY(hipEventRecord(w.last, s));
    w.used = true;
    w.last_stream = s;
    return BICOS_OK;
}

/* ------------------------------------------------------------- validation */

bool precision_known(int p) {
    return p == BICOS_PRECISION_SINGLE || p == BICOS_PRECISION_DOUBLE ||
           p == BICOS_PRECISION_HALF_MFMA || p == BICOS_PRECISION_MOMENTS ||
           p == BICOS_PRECISION_MOMENTS_FAST;
}

int elem_size(int type) {
    switch (type) {
        case BICOS_8U: return 1;
        case BICOS_16U: return 2;
        case BICOS_16S: return 2;
        case BICOS_32F: return 4;
        case BICOS_64F: return 8;
        default: return 0;
    }
}

int required_bits(uint32_t n, int mode) {
    // src/impl/cpu.cpp:118-120
    return mode == BICOS_MODE_FULL ? (int)(n * n - 2 * n + 3) : (int)(4 * n - 7);
}

int words_for_bits(int bits) {
    // src/impl/cpu.cpp:125-152
    if (bits <= 32) return 1;
    if (bits <= 64) return 2;
    if (bits <= 128) return 4;
    if (bits <= 256) return 8;
    return 0;
}

int check_stack(const bicos_image* s, uint32_t n, StackInfo* info, FrameTable* ft) {
    if (!s) return fail(BICOS_ERR_EXCEPTION, "null image stack");
    if (n < 2) return fail(BICOS_ERR_EXCEPTION, "need at least two images");  // cpu.cpp:106
    if (n > (uint32_t)MAX_FRAMES)
        return fail(BICOS_ERR_EXCEPTION, "Bad number of images");  // cuda.cu.in:105
    const int type = s[0].type;
    if (type != BICOS_8U && type != BICOS_16U)
        return fail(BICOS_ERR_EXCEPTION,
                    "bad input depths, only CV_8UC1 and CV_16UC1 are supported");  // cpu.cpp:109
    memset(ft, 0, sizeof(*ft));
    for (uint32_t t = 0; t < n; ++t) {
        if (s[t].type != type || s[t].rows != s[0].rows || s[t].cols != s[0].cols)
            return fail(BICOS_ERR_EXCEPTION, "all images of a stack must share size and type");
        if (!s[t].data || s[t].step < (int64_t)s[t].cols * elem_size(type))
            return fail(BICOS_ERR_EXCEPTION, "bad image data pointer or step");
        ft->ptr[t] = (uint64_t)s[t].data;
        ft->step[t] = s[t].step;
    }
    info->rows = s[0].rows;
    info->cols = s[0].cols;
    info->depth = type;
    if (info->rows < 1 || info->cols < 1) return fail(BICOS_ERR_EXCEPTION, "empty images");
    if (info->cols > 32767)
        return fail(BICOS_ERR_EXCEPTION, "image width above 32767 does not fit int16 disparities");
    return BICOS_OK;
}

// Frames one (n, rows, cols) allocation would have: equal row steps and a
// constant distance between consecutive frames.
bool uniformly_spaced(const FrameTable& f, uint32_t n) {
    const int64_t fs = n > 1 ? (int64_t)(f.ptr[1] - f.ptr[0]) : 0;
    for (uint32_t t = 1; t < n; ++t)
        if (f.step[t] != f.step[0] || (int64_t)(f.ptr[t] - f.ptr[0]) != (int64_t)t * fs) return false;
    return true;
}

size_t packed_bytes(const FrameTable& f, uint32_t n, const StackInfo& i) {
    return uniformly_spaced(f, n) ? 0 : (size_t)n * i.rows * i.cols * elem_size(i.depth);
}

// The strided view of a stack: in place, or packed into `scratch` (which
// must hold packed_bytes()) by one gather launch.
int make_stack(const FrameTable& f, uint32_t n, const StackInfo& i, void* scratch, Stack* out,
               hipStream_t s);

/* --------------------------------------------------------------- dispatch */

// 4-pixel-per-lane transform kernels (else one pixel per lane)
bool g_transform_vec = true;

bool rows_aligned(const Stack& f, size_t align) {
    return (uint64_t)f.base % align == 0 && (uint64_t)f.fstride % align == 0 &&
           (uint64_t)f.step % align == 0;
}

// Transform of one or two views (f1/out1 null: one view). The 4-pixel
// vectorised kernels need rows aligned to 4 samples; otherwise the
// one-pixel-per-lane kernels run.
template<typename TIn>
int run_transform(const Stack& f0, const Stack* f1, uint32_t n, int mode, int W, int H,
                  int words, uint32_t* out0, uint32_t* out1, hipStream_t s) {
    const int views = f1 ? 2 : 1;
    const bool vec = g_transform_vec && rows_aligned(f0, 4 * sizeof(TIn)) &&
                     (!f1 || rows_aligned(*f1, 4 * sizeof(TIn)));
    const Stack& g1 = f1 ? *f1 : f0;
    if (vec) {
        const dim3 grid((W + 256 * TPX - 1) / (256 * TPX), H, views), block(256);
        if (mode == BICOS_MODE_LIMITED) {
            // the deepest stack of each width (and 64 frames) with every
            // frame guard resolved at compile time
#define TL(WD, NM, PXV, NEV, G) LAUNCH("transform", (k_transform_limited_v4<TIn, WD, NM, PXV, NEV>), G, block, 0, s, f0, g1, (int)n, W, words, out0, out1)
            const dim3 grid2((W + 511) / 512, H, views);
            switch (words_for_bits(required_bits(n, mode))) {
                case 1: if (n == 9) TL(1, 9, TPX, 9, grid); else TL(1, 9, TPX, 0, grid); break;
                case 2: if (n == 17) TL(2, 17, TPX, 17, grid); else TL(2, 17, TPX, 0, grid); break;
                case 4: if (n == 33) TL(4, 33, TPX, 33, grid); else TL(4, 33, TPX, 0, grid); break;
                case 8:
                    if constexpr (sizeof(TIn) == 2) {  // 2 pixels per lane (register budget of 65 frames)
                        if (n == 64) TL(8, 65, 2, 64, grid2);
                        else if (n == 65) TL(8, 65, 2, 65, grid2);
                        else TL(8, 65, 2, 0, grid2);
                    } else {
                        TL(8, 65, TPX, 0, grid);
                    }
                    break;
                default: return fail(BICOS_ERR_INVALID_ARGUMENT, "unsupported stack depth");
            }
#undef TL
        } else {
            switch (n) {
#define TF(N) case N: LAUNCH("transform", (k_transform_full_v4<TIn, N>), grid, block, 0, s, f0, g1, W, words, out0, out1); break;
                TF(2) TF(3) TF(4) TF(5) TF(6) TF(7) TF(8) TF(9) TF(10) TF(11) TF(12) TF(13) TF(14) TF(15) TF(16)
#undef TF
                default: return fail(BICOS_ERR_INVALID_ARGUMENT, "unsupported stack depth");
            }
        }
        return BICOS_OK;
    }
    const dim3 grid((W + 255) / 256, H), block(256);
    for (int v = 0; v < views; ++v) {
        const Stack& f = v ? g1 : f0;
        uint32_t* out = v ? out1 : out0;
        if (mode == BICOS_MODE_LIMITED) {
            switch (words_for_bits(required_bits(n, mode))) {
                case 1: LAUNCH("transform_scalar", (k_transform_limited<TIn, 1, 9>), grid, block, 0, s, f, (int)n, W, words, out); break;
                case 2: LAUNCH("transform_scalar", (k_transform_limited<TIn, 2, 17>), grid, block, 0, s, f, (int)n, W, words, out); break;
                case 4: LAUNCH("transform_scalar", (k_transform_limited<TIn, 4, 33>), grid, block, 0, s, f, (int)n, W, words, out); break;
                case 8: LAUNCH("transform_scalar", (k_transform_limited<TIn, 8, 65>), grid, block, 0, s, f, (int)n, W, words, out); break;
                default: return fail(BICOS_ERR_INVALID_ARGUMENT, "unsupported stack depth");
            }
        } else {
            switch (n) {
#define TF(N) case N: LAUNCH("transform_scalar", (k_transform_full<TIn, N>), grid, block, 0, s, f, W, words, out); break;
                TF(2) TF(3) TF(4) TF(5) TF(6) TF(7) TF(8) TF(9) TF(10) TF(11) TF(12) TF(13) TF(14) TF(15) TF(16)
#undef TF
                default: return fail(BICOS_ERR_INVALID_ARGUMENT, "unsupported stack depth");
            }
        }
    }
    return BICOS_OK;
}

int transform_any(const Stack& f0, const Stack* f1, int depth, uint32_t n, int mode,
                  int W, int H, int words, uint32_t* out0, uint32_t* out1, hipStream_t s) {
    return depth == BICOS_8U ? run_transform<uint8_t>(f0, f1, n, mode, W, H, words, out0, out1, s)
                             : run_transform<uint16_t>(f0, f1, n, mode, W, H, words, out0, out1, s);
}

constexpr size_t SEARCH_LDS_BYTES = 64 * 1024;

template<int WORDS, int R>
int run_search_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes, int mode,
                 const int16_t* rev, int max_lr, int16_t* out, int64_t ostep, hipStream_t s) {
    const int chunk = (int)std::min<size_t>((size_t)W, SEARCH_LDS_BYTES / (WORDS * 4));
    const size_t lds = (size_t)chunk * WORDS * 4;
    const dim3 grid((W + SEARCH_THREADS * R - 1) / (SEARCH_THREADS * R), H), block(SEARCH_THREADS);
#define SK(ND, M) LAUNCH("search", (k_search<WORDS, ND, M, R>), grid, block, lds, s, dq, dc, W, chunk, rev, max_lr, out, ostep)
    if (nodupes) {
        if (mode == SEARCH_DISP) SK(true, SEARCH_DISP);
        else if (mode == SEARCH_REV) SK(true, SEARCH_REV);
        else SK(true, SEARCH_LR);
    } else {
        if (mode == SEARCH_DISP) SK(false, SEARCH_DISP);
        else if (mode == SEARCH_REV) SK(false, SEARCH_REV);
        else SK(false, SEARCH_LR);
    }
#undef SK
    return BICOS_OK;
}

// Search implementation: matrix cores (default) or vector ALU.
int g_search_impl = BICOS_SEARCH_MFMA;

// Candidate columns per LDS chunk of the full-row matrix-core searches,
// indexed by descriptor words (1, 2, 4, 8; 0 = no such kernel). fp4: up to
// 78 KiB of LDS (two blocks per CU); 256-bit descriptors take 512 columns
// (72 KiB), one prefetched column per each of the 512 threads, which keeps
// the kernel at four waves per SIMD (a 4096-column row is eight chunks). int8: chunks filling 64 KiB (measured best of the tried shapes for
// 1920-wide rows of 128-bit descriptors).
constexpr int FP4_CHUNK_COLS[9] = {0, 0, 1664, 0, 960, 0, 0, 0, 512};
constexpr int I8_CHUNK_COLS[9] = {0, 1344, 800, 0, 448, 0, 0, 0, 224};

template<int WORDS, int NB, int WAVES, int CH>
int run_search_mfma_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes,
                      int mode, const int16_t* rev, int max_lr, int16_t* out, int64_t ostep,
                      hipStream_t s) {
    constexpr size_t lds = (size_t)CH * (2 * WORDS + 1) * 16;
    static_assert(lds <= 64 * 1024, "LDS chunk too large");
    const dim3 grid((W + WAVES * NB * 32 - 1) / (WAVES * NB * 32), H), block(WAVES * 64);
    // keys straight from the matrix product where the column fits its field
#define SM(ND, M, E) LAUNCH("search_mfma", (k_search_mfma<WORDS, NB, ND, M, WAVES, CH, false, E>), grid, block, lds, s, dq, dc, W, rev, max_lr, out, ostep, (const int16_t*)nullptr, (const int*)nullptr, (int16_t*)nullptr)
#define SM2(ND, M) do { if (W <= EMBED_MAX_W) SM(ND, M, true); else SM(ND, M, false); } while (0)
    if (nodupes) {
        if (mode == SEARCH_DISP) SM2(true, SEARCH_DISP);
        else if (mode == SEARCH_REV) SM2(true, SEARCH_REV);
        else SM2(true, SEARCH_LR);
    } else {
        if (mode == SEARCH_DISP) SM2(false, SEARCH_DISP);
        else if (mode == SEARCH_REV) SM2(false, SEARCH_REV);
        else SM2(false, SEARCH_LR);
    }
#undef SM2
#undef SM
    return BICOS_OK;
}

// Dynamic LDS above the default 64 KiB per workgroup needs the function's
// attribute raised once.
int allow_dynamic_lds(const void* kernel, size_t lds) {
    static std::mutex mu;
    static std::map<const void*, size_t> raised;
    if (lds <= 65536) return BICOS_OK;
    std::lock_guard<std::mutex> lk(mu);
    if (raised[kernel] >= lds) return BICOS_OK;
    HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    raised[kernel] = lds;
    return BICOS_OK;
}

template<int WORDS, int NB, int WAVES, int CH>
int run_search_mfma4_w(const uint32_t* dq, const uint32_t* dc, int H, int W, bool nodupes,
                       int mode, const int16_t* rev, int max_lr, int16_t* out, int64_t ostep,
                       hipStream_t s) {
    constexpr size_t lds = (size_t)CH * (WORDS + 1) * 16;
    static_assert(2 * lds <= 160 * 1024, "two blocks per CU must fit the LDS");
    const dim3 grid((W + WAVES * NB * 32 - 1) / (WAVES * NB * 32), H), block(WAVES * 64);
#define SM(ND, M) do { \
        int rc_ = allow_dynamic_lds((const void*)k_search_mfma_fp4<WORDS, NB, ND, M, WAVES, CH>, lds); \
        if (rc_) return rc_; \
        LAUNCH("search_mfma", (k_search_mfma_fp4<WORDS, NB, ND, M, WAVES, CH>), grid, block, lds, s, dq, dc, W, rev, max_lr, out, ostep, (const int16_t*)nullptr, (const int*)nullptr, (int16_t*)nullptr); \
    } while (0)
    if (nodupes) {
        if (mode == SEARCH_DISP) SM(true, SEARCH_DISP);
        else if (mode == SEARCH_REV) SM(true, SEARCH_REV);
        else SM(true, SEARCH_LR);
    } else {
        if (mode == SEARCH_DISP) SM(false, SEARCH_DISP);
        else if (mode == SEARCH_REV) SM(false, SEARCH_REV);
        else SM(false, SEARCH_LR);
    }
#undef SM
    return BICOS_OK;
}

int run_search(const uint32_t* dq, const uint32_t* dc, int H, int W, int words, bool nodupes,
               int mode, const int16_t* rev, int max_lr, int16_t* out, int64_t ostep,
               hipStream_t s) {
    if (g_search_impl == BICOS_SEARCH_MFMA && words >= 2 && W <= EMBED_MAX_W) {
        switch (words) {
            // fp4 products (twice the int8 rate)
            case 2: return run_search_mfma4_w<2, 2, 8, FP4_CHUNK_COLS[2]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            case 4: return run_search_mfma4_w<4, 2, 8, FP4_CHUNK_COLS[4]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            case 8: return run_search_mfma4_w<8, 2, 8, FP4_CHUNK_COLS[8]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
        }
    }
    if (g_search_impl == BICOS_SEARCH_MFMA || g_search_impl == BICOS_SEARCH_MFMA_I8) {
        switch (words) {
            // 2 query tiles per wave, 8 waves per block
            case 1: return run_search_mfma_w<1, 2, 8, I8_CHUNK_COLS[1]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            case 2: return run_search_mfma_w<2, 2, 8, I8_CHUNK_COLS[2]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            case 4: return run_search_mfma_w<4, 2, 8, I8_CHUNK_COLS[4]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            case 8: return run_search_mfma_w<8, 2, 8, I8_CHUNK_COLS[8]>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
            default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
        }
    }
    switch (words) {
        case 1: return run_search_w<1, 16>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
        case 2: return run_search_w<2, 8>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
        case 4: return run_search_w<4, 8>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
        case 8: return run_search_w<8, 4>(dq, dc, H, W, nodupes, mode, rev, max_lr, out, ostep, s);
        default: return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
    }
}

// flags: bit 0 NODUPES, bit 1 CONSISTENCY (include/impl/common.hpp:46-47)
int search_flags(const uint32_t* d0, const uint32_t* d1, int H, int W, int words, int flags,
                 int max_lr, int16_t* out, int64_t ostep, int16_t* rev_scratch, hipStream_t s) {
    const bool nodupes = flags & 1;
    if (flags & 2) {
        int rc = run_search(d1, d0, H, W, words, nodupes, SEARCH_REV, nullptr, 0, rev_scratch, W, s);
        if (rc) return rc;
        return run_search(d0, d1, H, W, words, nodupes, SEARCH_LR, rev_scratch, max_lr, out, ostep, s);
    }
    return run_search(d0, d1, H, W, words, nodupes, SEARCH_DISP, nullptr, 0, out, ostep, s);
}

template<typename TIn, typename TPrec>
int run_agree_t(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                uint32_t n, int W, int H, float thr, int has_mv, float mv, TPrec* corr,
                int64_t cstep, hipStream_t s) {
    const dim3 grid((W + 255) / 256, H), block(256);
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        LAUNCH("agree", (k_agree<TIn, TPrec, decltype(NM)::value, decltype(NE)::value>), grid, block, 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, has_mv, mv, corr, cstep);
        return BICOS_OK;
    });
}

template<typename TIn, typename TPrec>
int run_subpixel_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                   uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                   float* out, int64_t ostep, TPrec* corr, int64_t cstep, hipStream_t s) {
    const dim3 grid((W + 255) / 256, H), block(256);
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        LAUNCH("agree_subpixel", (k_agree_subpixel<TIn, TPrec, decltype(NM)::value, decltype(NE)::value>), grid, block, 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, has_mv, mv, out, ostep, corr, cstep);
        return BICOS_OK;
    });
}

// HALF_MFMA concerns the subpixel candidates only; the single-point
// agreement is the exact FP32 one.
int run_agree(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1, int depth,
              uint32_t n, int W, int H, float thr, int has_mv, float mv_n, int precision,
              void* corr, int64_t cstep, hipStream_t s) {
    if (precision == BICOS_PRECISION_DOUBLE) {
        return depth == BICOS_8U
            ? run_agree_t<uint8_t, double>(disp, dstep, f0, f1, n, W, H, thr, has_mv, mv_n, (double*)corr, cstep, s)
            : run_agree_t<uint16_t, double>(disp, dstep, f0, f1, n, W, H, thr, has_mv, mv_n, (double*)corr, cstep, s);
    }
    return depth == BICOS_8U
        ? run_agree_t<uint8_t, float>(disp, dstep, f0, f1, n, W, H, thr, has_mv, mv_n, (float*)corr, cstep, s)
        : run_agree_t<uint16_t, float>(disp, dstep, f0, f1, n, W, H, thr, has_mv, mv_n, (float*)corr, cstep, s);
}

template<typename TIn>
int run_subpixel_half_t(const int16_t* disp, int64_t dstep, const Stack& f0,
                        const Stack& f1, uint32_t n, int W, int H, float thr, float step,
                        int has_mv, float mv, float* out, int64_t ostep, float* corr,
                        int64_t cstep, hipStream_t s) {
    const dim3 grid((W + 63) / 64, H), block(64);
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        constexpr int NMAX = decltype(NM)::value;
        constexpr size_t lds = sizeof(HalfLds<(NMAX + 15) / 16 * 16>);
        LAUNCH("agree_subpixel_half", (k_agree_subpixel_half<TIn, NMAX, decltype(NE)::value, 1>), grid, block, lds, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, has_mv, mv, out, ostep, corr, cstep);
        return BICOS_OK;
    });
}

// Subpixel kernel at Precision::SINGLE: pruned (default) or full scan.
int g_subpixel_impl = BICOS_SUBPIXEL_PRUNED;

template<typename TIn, typename TPrec>
int run_subpixel_pruned_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                          uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                          float* out, int64_t ostep, hipStream_t s) {
    const GridX gx = grid_x(step);
    const int npts = subpixel_points(step);
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        constexpr int NMAX = decltype(NM)::value, NEX = decltype(NE)::value;
        LAUNCH("agree_subpixel_pruned", (k_agree_subpixel_pruned<TIn, TPrec, NMAX, NEX>), dim3((W + 255) / 256, H), dim3(256), 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, npts, has_mv, mv, out, ostep, gx);
        return BICOS_OK;
    });
}

template<typename TIn>
int run_subpixel_fast_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                        uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                        float* out, int64_t ostep, hipStream_t s) {
    const GridX gx = grid_x(step);
    const int npts = subpixel_points(step);
    LAUNCH("agree_subpixel_fast", k_subpixel_fast<TIn>, dim3((W + 255) / 256, H), dim3(256), 0, s,
           disp, dstep, f0, f1, (int)n, W, thr, npts, gx, has_mv, mv, out, ostep);
    return BICOS_OK;
}

template<typename TIn>
int run_subpixel_fast_list_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                             uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                             float* out, int64_t ostep, const int16_t* list, const int* row_count,
                             hipStream_t s) {
    const GridX gx = grid_x(step);
    const int npts = subpixel_points(step);
    // up to four blocks per row (256 listed pixels each per pass)
    LAUNCH("agree_subpixel_fast_list", k_subpixel_fast_list<TIn>, dim3(H, 4), dim3(256), 0, s,
           disp, dstep, f0, f1, (int)n, W, thr, npts, gx, has_mv, mv, out, ostep, list, row_count);
    return BICOS_OK;
}

int run_subpixel_fast_list(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                           int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                           float mv_n, float* out, int64_t ostep, const int16_t* list,
                           const int* row_count, hipStream_t s) {
    return depth == BICOS_8U
        ? run_subpixel_fast_list_t<uint8_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, list, row_count, s)
        : run_subpixel_fast_list_t<uint16_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, list, row_count, s);
}

template<typename TIn>
int run_subpixel_moments_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                           uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                           float* out, int64_t ostep, uint8_t* status, hipStream_t s) {
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        constexpr int NMAX = decltype(NM)::value, NEX = decltype(NE)::value;
        LAUNCH("agree_subpixel_moments", (k_agree_subpixel_moments<TIn, NMAX, NEX>), dim3((W + 255) / 256, H), dim3(256), 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, subpixel_points(step), has_mv, mv, out, ostep, status);
        return BICOS_OK;
    });
}

template<typename TIn>
int run_moments_list_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                       uint32_t n, int W, int H, float thr, float step, int has_mv, float mv,
                       float* out, int64_t ostep, uint8_t* status, const int16_t* list,
                       const int* row_count, hipStream_t s) {
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        LAUNCH("agree_subpixel_moments_list", (k_agree_subpixel_moments_list<TIn, decltype(NM)::value, decltype(NE)::value>), dim3(LIST_BLOCKS), dim3(256), 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, subpixel_points(step), has_mv, mv, out, ostep, status, row_lists(list, row_count, W, H));
        return BICOS_OK;
    });
}

int run_moments_only(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                     int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                     float mv_n, float* out, int64_t ostep, const MomentsScratch* ms,
                     hipStream_t s) {
    return depth == BICOS_8U
        ? run_subpixel_moments_t<uint8_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, ms->status, s)
        : run_subpixel_moments_t<uint16_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, ms->status, s);
}

int run_moments_listed(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                       int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                       float mv_n, float* out, int64_t ostep, const MomentsScratch* ms,
                       const int16_t* list, const int* row_count, hipStream_t s) {
    return depth == BICOS_8U
        ? run_moments_list_t<uint8_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, ms->status, list, row_count, s)
        : run_moments_list_t<uint16_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, ms->status, list, row_count, s);
}

// row_start = exclusive prefix sum of row_count (one workgroup: each thread
// sums a contiguous chunk of rows, the chunk totals are scanned in LDS)
__global__ __launch_bounds__(1024) void k_row_scan(int* __restrict__ counts, int H) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (H + 1023) / 1024;
    const int r0 = min(t * chunk, H), r1 = min(r0 + chunk, H);
    int sum = 0;
    for (int r = r0; r < r1; ++r) sum += counts[r];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int* start = counts + H;
    int run = part[t] - sum;
    for (int r = r0; r < r1; ++r) {
        start[r] = run;
        run += counts[r];
    }
    if (t == 1023) start[H] = part[1023];
}

int run_row_scan(int* row_count, int H, hipStream_t s) {
    LAUNCH("row_scan", k_row_scan, dim3(1), dim3(1024), 0, s, row_count, H);
    return BICOS_OK;
}

int run_subpixel(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                 int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                 float mv_n, int precision, float* out, int64_t ostep, void* corr, int64_t cstep,
                 hipStream_t s, const MomentsScratch* ms) {
    if (precision == BICOS_PRECISION_MOMENTS && !corr) {
        // closed-form decisions, the exact stage on the pending pixels
        if (!ms) return fail(BICOS_ERR_INVALID_ARGUMENT, "MOMENTS needs its scratch buffers");
        return run_moments_only(disp, dstep, f0, f1, depth, n, W, H, thr, step, has_mv, mv_n, out,
                                ostep, ms, s);
    }
    if (precision == BICOS_PRECISION_MOMENTS_FAST && !corr) {
        const int npts = subpixel_points(step);
        if (npts >= 1 && npts <= 64)
            return depth == BICOS_8U
                ? run_subpixel_fast_t<uint8_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s)
                : run_subpixel_fast_t<uint16_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s);
    }
    if (precision == BICOS_PRECISION_MOMENTS || precision == BICOS_PRECISION_MOMENTS_FAST)
        precision = BICOS_PRECISION_SINGLE;  // a corrmap is written: the exact stage throughout
    if ((precision == BICOS_PRECISION_SINGLE || precision == BICOS_PRECISION_DOUBLE) && !corr &&
        g_subpixel_impl == BICOS_SUBPIXEL_PRUNED) {
        const int npts = subpixel_points(step);
        if (npts >= 1 && npts <= 64) {
            if (precision == BICOS_PRECISION_DOUBLE)
                return depth == BICOS_8U
                    ? run_subpixel_pruned_t<uint8_t, double>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s)
                    : run_subpixel_pruned_t<uint16_t, double>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s);
            return depth == BICOS_8U
                ? run_subpixel_pruned_t<uint8_t, float>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s)
                : run_subpixel_pruned_t<uint16_t, float>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, s);
        }
    }
    if (precision == BICOS_PRECISION_HALF_MFMA) {
        return depth == BICOS_8U
            ? run_subpixel_half_t<uint8_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, s)
            : run_subpixel_half_t<uint16_t>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, s);
    }
    if (precision == BICOS_PRECISION_DOUBLE) {
        return depth == BICOS_8U
            ? run_subpixel_t<uint8_t, double>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (double*)corr, cstep, s)
            : run_subpixel_t<uint16_t, double>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (double*)corr, cstep, s);
    }
    return depth == BICOS_8U
        ? run_subpixel_t<uint8_t, float>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, s)
        : run_subpixel_t<uint16_t, float>(disp, dstep, f0, f1, n, W, H, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, s);
}

template<typename TIn, typename TPrec>
int run_agree_list_t(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1, uint32_t n,
                     int W, float thr, int has_mv, float mv, TPrec* corr, int64_t cstep,
                     const int16_t* list, const int* row_count, int H, hipStream_t s) {
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        LAUNCH("agree_list", (k_agree_list<TIn, TPrec, decltype(NM)::value, decltype(NE)::value>), dim3(LIST_BLOCKS), dim3(256), 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, has_mv, mv, corr, cstep, row_lists(list, row_count, W, H));
        return BICOS_OK;
    });
}

template<typename TIn, typename TPrec>
int run_subpixel_list_t(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                        uint32_t n, int W, float thr, float step, int has_mv, float mv, float* out,
                        int64_t ostep, TPrec* corr, int64_t cstep, const int16_t* list,
                        const int* row_count, int H, hipStream_t s) {
    return with_nmax(n, [&](auto NM, auto NE) -> int {
        LAUNCH("agree_subpixel_list", (k_agree_subpixel_list<TIn, TPrec, decltype(NM)::value, decltype(NE)::value>), dim3(LIST_BLOCKS), dim3(256), 0, s,
               disp, dstep, f0, f1, (int)n, W, thr, step, has_mv, mv, out, ostep, corr, cstep, row_lists(list, row_count, W, H));
        return BICOS_OK;
    });
}

int run_agree_list(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1, int depth,
                   uint32_t n, int W, float thr, int has_mv, float mv_n, int precision, void* corr,
                   int64_t cstep, const int16_t* list, const int* row_count, int H,
                   hipStream_t s) {
    if (precision == BICOS_PRECISION_DOUBLE)
        return depth == BICOS_8U
            ? run_agree_list_t<uint8_t, double>(disp, dstep, f0, f1, n, W, thr, has_mv, mv_n, (double*)corr, cstep, list, row_count, H, s)
            : run_agree_list_t<uint16_t, double>(disp, dstep, f0, f1, n, W, thr, has_mv, mv_n, (double*)corr, cstep, list, row_count, H, s);
    return depth == BICOS_8U
        ? run_agree_list_t<uint8_t, float>(disp, dstep, f0, f1, n, W, thr, has_mv, mv_n, (float*)corr, cstep, list, row_count, H, s)
        : run_agree_list_t<uint16_t, float>(disp, dstep, f0, f1, n, W, thr, has_mv, mv_n, (float*)corr, cstep, list, row_count, H, s);
}

int run_subpixel_list(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                      int depth, uint32_t n, int W, float thr, float step, int has_mv, float mv_n,
                      int precision, float* out, int64_t ostep, void* corr, int64_t cstep,
                      const int16_t* list, const int* row_count, int H, hipStream_t s) {
    if (precision == BICOS_PRECISION_DOUBLE)
        return depth == BICOS_8U
            ? run_subpixel_list_t<uint8_t, double>(disp, dstep, f0, f1, n, W, thr, step, has_mv, mv_n, out, ostep, (double*)corr, cstep, list, row_count, H, s)
            : run_subpixel_list_t<uint16_t, double>(disp, dstep, f0, f1, n, W, thr, step, has_mv, mv_n, out, ostep, (double*)corr, cstep, list, row_count, H, s);
    return depth == BICOS_8U
        ? run_subpixel_list_t<uint8_t, float>(disp, dstep, f0, f1, n, W, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, list, row_count, H, s)
        : run_subpixel_list_t<uint16_t, float>(disp, dstep, f0, f1, n, W, thr, step, has_mv, mv_n, out, ostep, (float*)corr, cstep, list, row_count, H, s);
}

// Number of points of the subpixel grid `for (float x = -1; x <= 1; x += step)`
// (agree.hpp:162), counted with the same float arithmetic the kernel uses;
// -1 when the step is not positive or so small that the loop would not end
// (the reference would spin forever) or would exceed 65536 points.
int subpixel_points(float step) {
    if (!(step > 0.0f)) return -1;
    int k = 0;
    for (float x = -1.0f; x <= 1.0f; ++k) {
        const float nx = x + step;
        if (!(nx > x) || k > 65536) return -1;
        x = nx;
    }
    return k;
}

int check_step(float step) {
    if (subpixel_points(step) < 0)
        return fail(BICOS_ERR_INVALID_ARGUMENT,
                    "subpixel_step must be positive and give at most 65536 grid points");
    return BICOS_OK;
}

int check_out(const bicos_image* img, int rows, int cols, int type, const char* what) {
    if (!img || !img->data)
        return fail(BICOS_ERR_EXCEPTION, std::string(what) + ": null output image");
    if (img->rows != rows || img->cols != cols)
        return fail(BICOS_ERR_EXCEPTION, std::string(what) + ": size does not match the input");
    if (img->type != type)
        return fail(BICOS_ERR_EXCEPTION, std::string(what) + ": wrong element type");
    const int es = elem_size(type);
    if (img->step % es != 0 || img->step < (int64_t)cols * es)
        return fail(BICOS_ERR_EXCEPTION, std::string(what) + ": bad row step");
    return BICOS_OK;
}

int make_stack(const FrameTable& f, uint32_t n, const StackInfo& i, void* scratch, Stack* out,
               hipStream_t s) {
    if (uniformly_spaced(f, n)) {
        out->base = (const uint8_t*)f.ptr[0];
        out->fstride = n > 1 ? (int64_t)(f.ptr[1] - f.ptr[0]) : 0;
        out->step = f.step[0];
        return BICOS_OK;
    }
    const dim3 grid((i.cols + 1023) / 1024, i.rows, n), block(256);
    if (i.depth == BICOS_8U)
        LAUNCH("pack_frames", k_pack_frames<uint8_t>, grid, block, 0, s, f, i.cols, (uint8_t*)scratch);
    else
        LAUNCH("pack_frames", k_pack_frames<uint16_t>, grid, block, 0, s, f, i.cols, (uint16_t*)scratch);
    const int es = elem_size(i.depth);
    out->base = (const uint8_t*)scratch;
    out->fstride = (int64_t)i.rows * i.cols * es;
    out->step = (int64_t)i.cols * es;
    return BICOS_OK;
}

int current_device(int* dev) {
    HIP_TRY(hipGetDevice(dev));
    return BICOS_OK;
}

// Runs fn(stacks...) with strided views of up to two stacks, packing
// non-uniformly spaced ones into the workspace.
template<typename F>
int with_stacks(const FrameTable* f, const StackInfo* info, int count, uint32_t n, hipStream_t s,
                F&& fn, size_t extra = 0) {
    size_t pk[2] = {0, 0};
    for (int k = 0; k < count; ++k) pk[k] = align_up(packed_bytes(f[k], n, info[k]), 256);
    int dev, rc;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws = nullptr;
    if ((rc = ws_acquire(dev, pk[0] + pk[1] + extra + 256, s, &ws))) return rc;
    Stack v[2];
    for (int k = 0; k < count && !rc; ++k)
        rc = make_stack(f[k], n, info[k], (uint8_t*)ws + (k ? pk[0] : 0), &v[k], s);
    if (!rc) rc = fn(v, (uint8_t*)ws + pk[0] + pk[1]);
    int rc2 = ws_release(dev, s);
    return rc ? rc : rc2;
}

}  // namespace bicos_hip

namespace bicos_hip {
namespace {

/* ------------------------------------------------- no-allocation scratch */

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)96 << 20;  // 96 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;

// Scratch of bicos_hip_match_noalloc: a file-scope device array, one copy
// per GPU (held while the process lives). In host code the array's name is
// only a stand-in; its device address on each GPU comes from
// hipGetSymbolAddress with that GPU current, taken once per device and kept
// in g_noalloc_ptr. Calls on one GPU share it (overlapping calls race); no
// kernel of the plain path needs it cleared (every byte a call reads was
// written earlier in the same call).
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(int dev, void** out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(BICOS_ERR_HIP, "no device address for the no-allocation scratch");
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return BICOS_OK;
}

}  // namespace

// BICOS::match (src/impl/cuda.cu.in:233-292). The scratch (descriptor
// planes, the int16 disparity before the subpixel stage, the reverse-search
// columns, packed frames, the MOMENTS status) comes from the per-device
// workspace, or with noalloc from the static scratch array (no allocation;
// calls needing more than NOALLOC_SCRATCH_BYTES are rejected).
int match_common(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                 const bicos_config* cfg, bicos_image* disparity, bicos_image* corrmap,
                 void* stream, bool noalloc) {
    if (!cfg) return fail(BICOS_ERR_INVALID_ARGUMENT, "null config");
    hipStream_t s = (hipStream_t)stream;
    StackInfo i0, i1;
    FrameTable f0, f1;
    int rc;
    if ((rc = check_stack(stack0, n, &i0, &f0))) return rc;
    if ((rc = check_stack(stack1, n, &i1, &f1))) return rc;
    if (i0.rows != i1.rows || i0.cols != i1.cols || i0.depth != i1.depth)
        return fail(BICOS_ERR_EXCEPTION, "the two stacks differ in size or type");
    if (cfg->mode != BICOS_MODE_LIMITED && cfg->mode != BICOS_MODE_FULL)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown transform mode");
    if (cfg->variant != BICOS_VARIANT_NO_DUPLICATES && cfg->variant != BICOS_VARIANT_CONSISTENCY)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented");  // cuda.cu.in:51
    if (!precision_known(cfg->precision))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented precision");

    const int bits = required_bits(n, cfg->mode);
    const int words = words_for_bits(bits);
    if (!words)
        return fail(BICOS_ERR_INVALID_ARGUMENT,
                    "input stacks too large, would require " + std::to_string(bits) + " bits");
    const bool thr = cfg->has_nxcorr_threshold != 0;
    const bool subpix = thr && cfg->has_subpixel_step;
    if (subpix && (rc = check_step(cfg->subpixel_step))) return rc;

    const int H = i0.rows, W = i0.cols;
    if ((rc = check_out(disparity, H, W, bicos_hip_disparity_type(cfg), "disparity"))) return rc;
    if (corrmap && thr && (rc = check_out(corrmap, H, W, bicos_hip_corrmap_type(cfg), "corrmap")))
        return rc;

    const bool consistency = cfg->variant == BICOS_VARIANT_CONSISTENCY;
    const size_t npx = (size_t)W * H;
    const size_t desc_bytes = align_up(npx * words * 4, 256);
    const size_t raw_bytes = subpix ? align_up(npx * 2, 256) : 0;
    const size_t rev_bytes = consistency ? align_up(npx * 2, 256) : 0;
    const size_t pk0 = align_up(packed_bytes(f0, n, i0), 256), pk1 = align_up(packed_bytes(f1, n, i1), 256);
    const bool moments = subpix && cfg->precision == BICOS_PRECISION_MOMENTS;
    const size_t mom_bytes = moments ? MomentsScratch::bytes(npx) : 0;

    const size_t need = 2 * desc_bytes + raw_bytes + rev_bytes + pk0 + pk1 + mom_bytes;
    int dev;
    if ((rc = current_device(&dev))) return rc;
    std::unique_lock<std::mutex> lk(g_ws_mu, std::defer_lock);
    void* ws;
    if (noalloc) {
        if (need > NOALLOC_SCRATCH_BYTES)
            return fail(BICOS_ERR_INVALID_ARGUMENT,
                        "this match needs " + std::to_string(need) + " bytes of scratch; the "
                        "no-allocation entry point has " + std::to_string(NOALLOC_SCRATCH_BYTES));
        if ((rc = noalloc_scratch(dev, &ws))) return rc;
    } else {
        lk.lock();
        if ((rc = ws_acquire(dev, need, s, &ws))) return rc;
    }
    uint8_t* p = (uint8_t*)ws;
    uint32_t* desc0 = (uint32_t*)p;
    uint32_t* desc1 = (uint32_t*)(p + desc_bytes);
    int16_t* raw = subpix ? (int16_t*)(p + 2 * desc_bytes) : nullptr;
    int16_t* rev = consistency ? (int16_t*)(p + 2 * desc_bytes + raw_bytes) : nullptr;
    uint8_t* pack0 = p + 2 * desc_bytes + raw_bytes + rev_bytes;
    uint8_t* pack1 = pack0 + pk0;
    const MomentsScratch ms = MomentsScratch::carve(pack1 + pk1);

    auto body = [&]() -> int {
        int r;
        Stack v0, v1;
        if ((r = make_stack(f0, n, i0, pack0, &v0, s))) return r;
        if ((r = make_stack(f1, n, i1, pack1, &v1, s))) return r;
        if ((r = transform_any(v0, &v1, i0.depth, n, cfg->mode, W, H, words, desc0, desc1, s)))
            return r;

        int flags = 1, lr = -1;  // cpu.cpp:70-77
        if (consistency) {
            flags = 2 | (cfg->no_dupes ? 1 : 0);
            lr = cfg->max_lr_diff;
        }
        int16_t* d16 = subpix ? raw : (int16_t*)disparity->data;
        const int64_t d16_step = subpix ? W : disparity->step / 2;
        if ((r = search_flags(desc0, desc1, H, W, words, flags, lr, d16, d16_step, rev, s))) return r;
        if (!thr) return BICOS_OK;

        const float mv_n = cfg->has_min_variance ? cfg->min_variance * (float)n : 0.0f;  // cpu.cpp:123
        void* cm = (corrmap && thr) ? corrmap->data : nullptr;
        const int64_t cstep = cm ? corrmap->step / elem_size(corrmap->type) : 0;
        if (subpix)
            return run_subpixel(raw, W, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                                cfg->subpixel_step, cfg->has_min_variance, mv_n, cfg->precision,
                                (float*)disparity->data, disparity->step / 4, cm, cstep, s,
                                moments ? &ms : nullptr);
        return run_agree(d16, d16_step, v0, v1, i0.depth, n, W, H, cfg->nxcorr_threshold,
                         cfg->has_min_variance, mv_n, cfg->precision, cm, cstep, s);
    };
    rc = body();
    if (noalloc) return rc;
    int rc2 = ws_release(dev, s);
    return rc ? rc : rc2;
}

}  // namespace bicos_hip

using namespace bicos_hip;

/* ================================================================== C ABI */

extern "C" {

void bicos_hip_config_default(bicos_config* c) {
    // include/common.hpp:73-82, include/common.hpp:66-69
    memset(c, 0, sizeof(*c));
    c->has_nxcorr_threshold = 1;
    c->nxcorr_threshold = 0.5f;
    c->mode = BICOS_MODE_LIMITED;
    c->precision = BICOS_PRECISION_SINGLE;
    c->variant = BICOS_VARIANT_NO_DUPLICATES;
    c->max_lr_diff = 1;
    c->no_dupes = 0;
}

int bicos_hip_disparity_type(const bicos_config* c) {
    return (c->has_nxcorr_threshold && c->has_subpixel_step) ? BICOS_32F : BICOS_16S;
}

int bicos_hip_corrmap_type(const bicos_config* c) {
    if (!c->has_nxcorr_threshold) return -1;
    return c->precision == BICOS_PRECISION_DOUBLE ? BICOS_64F : BICOS_32F;
}

int bicos_hip_descriptor_bits(uint32_t n, int32_t mode) {
    if (n < 2) return -1;
    const int w = words_for_bits(required_bits(n, mode));
    return w ? 32 * w : -1;
}

const char* bicos_hip_last_error(void) { return g_error.c_str(); }

int bicos_hip_match(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                    const bicos_config* cfg, bicos_image* disparity, bicos_image* corrmap,
                    void* stream) {
    return match_common(stack0, stack1, n, cfg, disparity, corrmap, stream, false);
}

int bicos_hip_match_noalloc(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                            const bicos_config* cfg, bicos_image* disparity, bicos_image* corrmap,
                            void* stream) {
    return match_common(stack0, stack1, n, cfg, disparity, corrmap, stream, true);
}

size_t bicos_hip_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int bicos_hip_set_search_impl(int32_t impl) {
    if (impl != BICOS_SEARCH_MFMA && impl != BICOS_SEARCH_VALU && impl != BICOS_SEARCH_MFMA_I8)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown search implementation");
    g_search_impl = impl;
    return BICOS_OK;
}

int bicos_hip_search_chunk_cols(int32_t impl, int32_t words) {
    if (words != 1 && words != 2 && words != 4 && words != 8) return -1;
    if (impl == BICOS_SEARCH_MFMA && words >= 2) return FP4_CHUNK_COLS[words];
    if (impl == BICOS_SEARCH_MFMA || impl == BICOS_SEARCH_MFMA_I8) return I8_CHUNK_COLS[words];
    return impl == BICOS_SEARCH_VALU ? 0 : -1;
}

int bicos_hip_set_subpixel_impl(int32_t impl) {
    if (impl != BICOS_SUBPIXEL_PRUNED && impl != BICOS_SUBPIXEL_SCAN)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown subpixel implementation");
    g_subpixel_impl = impl;
    return BICOS_OK;
}

int bicos_hip_release_workspace(void) {
    int dev, rc;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (int slot = 0; slot < WS_SLOTS; ++slot) {
        auto it = g_ws.find(dev * WS_SLOTS + slot);
        if (it == g_ws.end()) continue;
        Workspace& w = it->second;
        if (w.used) HIP_TRY(hipEventSynchronize(w.last));
        if (w.buf) HIP_TRY(hipFree(w.buf));
        if (w.last) HIP_TRY(hipEventDestroy(w.last));
        g_ws.erase(it);
    }
    return BICOS_OK;
}

int bicos_hip_transform(const bicos_image* frames, uint32_t n, int32_t mode, int32_t words,
                        uint32_t* out, void* stream) {
    StackInfo info;
    FrameTable f;
    int rc;
    if ((rc = check_stack(frames, n, &info, &f))) return rc;
    if (mode != BICOS_MODE_LIMITED && mode != BICOS_MODE_FULL)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unknown transform mode");
    const int bits = required_bits(n, mode);
    const int natural = words_for_bits(bits);
    if (!natural)
        return fail(BICOS_ERR_INVALID_ARGUMENT,
                    "input stacks too large, would require " + std::to_string(bits) + " bits");
    if (words != 1 && words != 2 && words != 4 && words != 8)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor words must be 1, 2, 4 or 8");
    if (words < natural)
        return fail(BICOS_ERR_INVALID_ARGUMENT, "descriptor too narrow for this stack");
    if (!out) return fail(BICOS_ERR_EXCEPTION, "null output");
    hipStream_t s = (hipStream_t)stream;
    return with_stacks(&f, &info, 1, n, s, [&](const Stack* v, void*) {
        return transform_any(v[0], nullptr, info.depth, n, mode, info.cols, info.rows, words, out,
                             nullptr, s);
    });
}

int bicos_hip_search(const uint32_t* desc0, const uint32_t* desc1, int32_t rows, int32_t cols,
                     int32_t words, int32_t flags, int32_t max_lr_diff, int16_t* out,
                     void* stream) {
    if (!desc0 || !desc1 || !out) return fail(BICOS_ERR_EXCEPTION, "null buffer");
    if (rows < 1 || cols < 1 || cols > 32767) return fail(BICOS_ERR_EXCEPTION, "bad size");
    if (flags < 0 || flags > 3) return fail(BICOS_ERR_INVALID_ARGUMENT, "bad flags");
    hipStream_t s = (hipStream_t)stream;
    if (!(flags & 2))
        return search_flags(desc0, desc1, rows, cols, words, flags, max_lr_diff, out, cols,
                            nullptr, s);
    int dev, rc;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    void* ws;
    if ((rc = ws_acquire(dev, align_up((size_t)rows * cols * 2, 256), s, &ws))) return rc;
    rc = search_flags(desc0, desc1, rows, cols, words, flags, max_lr_diff, out, cols,
                      (int16_t*)ws, s);
    int rc2 = ws_release(dev, s);
    return rc ? rc : rc2;
}

int bicos_hip_agree(int16_t* disparity, const bicos_image* stack0, const bicos_image* stack1,
                    uint32_t n, float thr, int32_t has_min_variance, float min_variance,
                    int32_t precision, void* corrmap, void* stream) {
    StackInfo i[2];
    FrameTable f[2];
    int rc;
    if ((rc = check_stack(stack0, n, &i[0], &f[0]))) return rc;
    if ((rc = check_stack(stack1, n, &i[1], &f[1]))) return rc;
    if (!disparity) return fail(BICOS_ERR_EXCEPTION, "null disparity");
    if (!precision_known(precision))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented precision");
    const float mv_n = has_min_variance ? min_variance * (float)n : 0.0f;
    hipStream_t s = (hipStream_t)stream;
    return with_stacks(f, i, 2, n, s, [&](const Stack* v, void*) {
        return run_agree(disparity, i[0].cols, v[0], v[1], i[0].depth, n, i[0].cols, i[0].rows,
                         thr, has_min_variance, mv_n, precision, corrmap, i[0].cols, s);
    });
}

int bicos_hip_agree_subpixel(const int16_t* disparity, const bicos_image* stack0,
                             const bicos_image* stack1, uint32_t n, float thr,
                             float subpixel_step, int32_t has_min_variance, float min_variance,
                             int32_t precision, float* out, void* corrmap, void* stream) {
    StackInfo i[2];
    FrameTable f[2];
    int rc;
    if ((rc = check_stack(stack0, n, &i[0], &f[0]))) return rc;
    if ((rc = check_stack(stack1, n, &i[1], &f[1]))) return rc;
    if (!disparity || !out) return fail(BICOS_ERR_EXCEPTION, "null buffer");
    if ((rc = check_step(subpixel_step))) return rc;
    if (!precision_known(precision))
        return fail(BICOS_ERR_INVALID_ARGUMENT, "unimplemented precision");
    const float mv_n = has_min_variance ? min_variance * (float)n : 0.0f;
    hipStream_t s = (hipStream_t)stream;
    const size_t npx = (size_t)i[0].rows * i[0].cols;
    return with_stacks(f, i, 2, n, s, [&](const Stack* v, void* extra) {
        const MomentsScratch ms = MomentsScratch::carve(extra);
        return run_subpixel(disparity, i[0].cols, v[0], v[1], i[0].depth, n, i[0].cols, i[0].rows,
                            thr, subpixel_step, has_min_variance, mv_n, precision, out,
                            i[0].cols, corrmap, i[0].cols, s, &ms);
    }, precision == BICOS_PRECISION_MOMENTS ? MomentsScratch::bytes(npx) : 0);
}

int bicos_hip_timing_enable(int32_t enable) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    for (auto& t : g_timed) {
        (void)hipEventSynchronize(t.stop);
        g_event_pool.push_back(t.start);
        g_event_pool.push_back(t.stop);
    }
    g_timed.clear();
    g_timing = enable != 0;
    return BICOS_OK;
}

int bicos_hip_timing_read(bicos_kernel_stat* out, int32_t cap) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    std::vector<bicos_kernel_stat> stats;
    for (auto& t : g_timed) {
        HIP_TRY(hipEventSynchronize(t.stop));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t.start, t.stop));
        auto it = std::find_if(stats.begin(), stats.end(),
                               [&](const bicos_kernel_stat& k) { return strcmp(k.name, t.name) == 0; });
        if (it == stats.end()) {
            bicos_kernel_stat k{};
            strncpy(k.name, t.name, sizeof(k.name) - 1);
            stats.push_back(k);
            it = stats.end() - 1;
        }
        it->launches += 1;
        it->total_ms += ms;
    }
    for (int i = 0; i < (int)stats.size() && i < cap; ++i) out[i] = stats[i];
    return (int)stats.size();
}

void* bicos_hip_malloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        g_error = "hipMalloc failed";
        return nullptr;
    }
    return p;
}

int bicos_hip_free(void* p) {
    HIP_TRY(hipFree(p));
    return BICOS_OK;
}

int bicos_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BICOS_OK;
}

int bicos_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BICOS_OK;
}

int bicos_hip_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return BICOS_OK;
}

}  // extern "C"
