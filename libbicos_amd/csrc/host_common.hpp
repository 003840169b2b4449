// Host-side helpers shared by the translation units of libbicos_hip.so:
// error reporting, timed launches, the per-device workspace, input
// validation and the stage dispatchers of the plain matching path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bicos_hip.h"
#include "device_common.hpp"

namespace bicos_hip {

extern thread_local std::string g_error;
int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(BICOS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

/* ----------------------------------------------------------------- timing */

struct TimedLaunch {
    const char* name;
    hipEvent_t start, stop;
};

extern std::mutex g_timing_mu;
extern bool g_timing;
extern std::vector<TimedLaunch> g_timed;
hipEvent_t pooled_event();

template<typename Kernel, typename... Args>
int launch(const char* name, Kernel kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s,
           Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_timing) {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        e0 = pooled_event();
        e1 = pooled_event();
        if (e0) (void)hipEventRecord(e0, s);
    }
    hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return fail(BICOS_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(err));
    if (e0 && e1) {
        (void)hipEventRecord(e1, s);
        std::lock_guard<std::mutex> lk(g_timing_mu);
        g_timed.push_back({name, e0, e1});
    }
    return BICOS_OK;
}

#define LAUNCH(...)                                \
    do {                                           \
        int rc_ = launch(__VA_ARGS__);             \
        if (rc_ != BICOS_OK) return rc_;           \
    } while (0)

/* ------------------------------------------------ workspace, validation */

// Device buffers kept between calls, one per (device, slot); guarded by
// g_ws_mu. ws_acquire returns a buffer of at least `bytes` that is free for
// use on `s` (the stream waits for the last call that used it); ws_release
// marks the end of that use on `s`.
enum WorkspaceSlot : int { WS_MATCH = 0, WS_SCENE = 1, WS_WINDOW = 2, WS_POINTS = 3, WS_SLOTS = 4 };
extern std::mutex g_ws_mu;
int ws_acquire(int dev, size_t bytes, hipStream_t s, void** out, int slot = WS_MATCH);
int ws_release(int dev, hipStream_t s, int slot = WS_MATCH);
int current_device(int* dev);

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct StackInfo {
    int rows = 0, cols = 0, depth = 0;
};

int elem_size(int type);
bool precision_known(int precision);
int required_bits(uint32_t n, int mode);
int words_for_bits(int bits);
int check_stack(const bicos_image* s, uint32_t n, StackInfo* info, FrameTable* ft);
bool uniformly_spaced(const FrameTable& f, uint32_t n);
size_t packed_bytes(const FrameTable& f, uint32_t n, const StackInfo& i);
int make_stack(const FrameTable& f, uint32_t n, const StackInfo& i, void* scratch, Stack* out,
               hipStream_t s);
int check_step(float step);
// Raises a kernel's dynamic LDS limit once where `lds` exceeds the default 64 KiB
int allow_dynamic_lds(const void* kernel, size_t lds);
int subpixel_points(float step);

// stack depth buckets of the NCC kernels (registers per lane scale with NMAX)
template<typename F>
int with_nmax(uint32_t n, F&& f) {
    if (n == 12) return f(std::integral_constant<int, 12>{}, std::integral_constant<int, 12>{});
    if (n == 33) return f(std::integral_constant<int, 33>{}, std::integral_constant<int, 33>{});
    if (n == 64) return f(std::integral_constant<int, 64>{}, std::integral_constant<int, 64>{});
    if (n <= 8) return f(std::integral_constant<int, 8>{}, std::integral_constant<int, 0>{});
    if (n <= 16) return f(std::integral_constant<int, 16>{}, std::integral_constant<int, 0>{});
    if (n <= 24) return f(std::integral_constant<int, 24>{}, std::integral_constant<int, 0>{});
    if (n <= 33) return f(std::integral_constant<int, 33>{}, std::integral_constant<int, 0>{});
    if (n <= 48) return f(std::integral_constant<int, 48>{}, std::integral_constant<int, 0>{});
    return f(std::integral_constant<int, 65>{}, std::integral_constant<int, 0>{});
}

// The grid points x_k of the subpixel scan as the reference's loop forms them
// (x_0 = -1, x_k = x_(k-1) + step in FP32, agree.hpp:160); the first 64.
inline GridX grid_x(float step) {
    GridX gx;
    const int npts = subpixel_points(step);
    float x = -1.0f;
    for (int k = 0; k < 64; ++k) {
        gx.x[k] = k < npts ? x : 0.0f;
        x += step;
    }
    return gx;
}

// MOMENTS_FAST subpixel stage of the listed pixels (row r: row_count[r]
// columns at list + r * W; needs no row_start)
int run_subpixel_fast_list(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                           int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                           float mv_n, float* out, int64_t ostep, const int16_t* list,
                           const int* row_count, hipStream_t s);
int check_out(const bicos_image* img, int rows, int cols, int type, const char* what);

/* ------------------------------------------------------ stage dispatch */

// Row-search implementation (BICOS_SEARCH_MFMA or BICOS_SEARCH_VALU).
extern int g_search_impl;
// Subpixel kernel at Precision::SINGLE (BICOS_SUBPIXEL_PRUNED or _SCAN).
extern int g_subpixel_impl;

int transform_any(const Stack& f0, const Stack* f1, int depth, uint32_t n, int mode,
                  int W, int H, int words, uint32_t* out0, uint32_t* out1, hipStream_t s);
int search_flags(const uint32_t* d0, const uint32_t* d1, int H, int W, int words, int flags,
                 int max_lr, int16_t* out, int64_t ostep, int16_t* rev_scratch, hipStream_t s);
int run_agree(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1, int depth,
              uint32_t n, int W, int H, float thr, int has_mv, float mv_n, int precision,
              void* corr, int64_t cstep, hipStream_t s);
// Per-row pixel lists keep their counts in one int buffer of 2 * rows + 1
// values: row_count[r] (rows values), then row_start, the exclusive prefix
// sum (rows + 1 values) that numbers every entry densely; the list kernels
// walk that numbering with a fixed grid, so one crowded row does not set
// their duration.
inline size_t list_counts_bytes(int rows) { return align_up((size_t)(2 * rows + 1) * 4, 256); }
inline RowLists row_lists(const int16_t* list, const int* row_count, int W, int H) {
    return RowLists{list, row_count, row_count + H, W, H};
}
// row_count + H = prefix sums of row_count[0 .. H)
int run_row_scan(int* row_count, int H, hipStream_t s);
// workgroups of the list kernels (grid-stride over the dense numbering)
constexpr int LIST_BLOCKS = 2048;

// Scratch of the MOMENTS subpixel stage: per-pixel status (ST_HOPELESS:
// the left variance rejects every match; read by the fallback rule)
struct MomentsScratch {
    uint8_t* status;  // rows * cols
    static size_t bytes(size_t npx) { return align_up(npx, 256); }
    static MomentsScratch carve(void* p) { return MomentsScratch{(uint8_t*)p}; }
};

// MOMENTS subpixel stage of every pixel (closed form, exact stage on the
// pixels it leaves pending), writing the status
int run_moments_only(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                     int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                     float mv_n, float* out, int64_t ostep, const MomentsScratch* ms,
                     hipStream_t s);

// the same for listed pixels only (lists with row_start built)
int run_moments_listed(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                       int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                       float mv_n, float* out, int64_t ostep, const MomentsScratch* ms,
                       const int16_t* list, const int* row_count, hipStream_t s);

// ms: required for PRECISION_MOMENTS without a corrmap (with a corrmap the
// exact FP32 stage runs)
int run_subpixel(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                 int depth, uint32_t n, int W, int H, float thr, float step, int has_mv,
                 float mv_n, int precision, float* out, int64_t ostep, void* corr, int64_t cstep,
                 hipStream_t s, const MomentsScratch* ms = nullptr);
// NCC stage of listed pixels only (row r: row_count[r] columns at
// list + r * W): exact FP32 arithmetic, FP64 for PRECISION_DOUBLE.
int run_agree_list(int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1, int depth,
                   uint32_t n, int W, float thr, int has_mv, float mv_n, int precision, void* corr,
                   int64_t cstep, const int16_t* list, const int* row_count, int H,
                   hipStream_t s);
int run_subpixel_list(const int16_t* disp, int64_t dstep, const Stack& f0, const Stack& f1,
                      int depth, uint32_t n, int W, float thr, float step, int has_mv, float mv_n,
                      int precision, float* out, int64_t ostep, void* corr, int64_t cstep,
                      const int16_t* list, const int* row_count, int H, hipStream_t s);

}  // namespace bicos_hip
