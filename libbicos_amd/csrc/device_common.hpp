// Shared device-side definitions of the BICOS gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bicos_hip {

// A view's frame planes, passed by value in the kernel argument segment:
// frame t of the view starts at ptr[t], its rows are step[t] bytes apart.
// 65 frames is the deepest stack any descriptor width admits
// (limited mode, 256-bit descriptors; include/impl/common.hpp:61-67).
constexpr int MAX_FRAMES = 65;

struct FrameTable {
    uint64_t ptr[MAX_FRAMES];
    int64_t step[MAX_FRAMES];
};

// The frames of one view as the compute kernels see them: frame t, row r
// starts at base + t * fstride + r * step (bytes). Uniformly spaced frames
// (one (n, rows, cols) allocation) are used in place; others are packed
// into this form first (k_pack_frames).
struct Stack {
    const uint8_t* base;
    int64_t fstride;
    int64_t step;
};

constexpr int16_t INVALID16 = -32768;  // INVALID_DISP<int16_t> (include/common.hpp:34-37)

// A listed pixel's full-row match replaces its windowed match in dst. When
// the two are the same valid match, the NCC stage has already scored it
// (the pixel is listed because that match was rejected): LISTED_SAME16 is
// stored instead, and the listed NCC stage leaves the pixel's outputs as the
// windowed pass wrote them. (-32767 is no disparity: |d| < 32767 for rows of
// at most 32767 columns.)
constexpr int16_t LISTED_SAME16 = -32767;
__device__ __forceinline__ void set_listed(int16_t* dst, int16_t v) {
    *dst = (v == *dst && v != INVALID16) ? LISTED_SAME16 : v;
}

// The grid points x_k of the subpixel scan (x_0 = -1, x_k = x_(k-1) + step
// in FP32, agree.hpp:160), formed on the host the way the reference's loop
// forms them (grid_x); passed by value where npts <= 64.
struct GridX {
    float x[64];
};

template<typename T>
__device__ __forceinline__ const T* frame_row(const Stack& f, int t, int row) {
    return (const T*)(f.base + (int64_t)t * f.fstride + (int64_t)row * f.step);
}

template<typename T>
__device__ __forceinline__ uint32_t load_sample_cached(const Stack& f, int t, int row, int col) {
    return (uint32_t)frame_row<T>(f, t, row)[col];
}

// gathers n frame planes given by a pointer table into one dense stack
// (frame stride rows * cols, row step cols), 4 samples per lane
template<typename T>
__global__ __launch_bounds__(256) void k_pack_frames(const FrameTable f, int W, T* __restrict__ out) {
    const int col = (blockIdx.x * 256 + threadIdx.x) * 4, row = blockIdx.y, t = blockIdx.z;
    if (col >= W)
        return;
    const T* src = (const T*)(f.ptr[t] + (uint64_t)row * (uint64_t)f.step[t]);
    T* dst = out + ((size_t)t * gridDim.y + row) * W;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (col + k < W) dst[col + k] = src[col + k];
}

// Per-row pixel lists: row r holds row_count[r] columns at cols + r * W (any
// order); row_start, when built (k_row_scan), is the exclusive prefix sum of
// row_count (rows + 1 values), which numbers all entries densely.
struct RowLists {
    const int16_t* cols;
    const int* row_count;
    const int* row_start;
    int W, H;
};

// entry i (< row_start[H]) of the dense numbering: its row and column
__device__ __forceinline__ void list_entry(const RowLists& l, int i, int& row, int& col) {
    int lo = 0, hi = l.H - 1;  // largest r with row_start[r] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (l.row_start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    row = lo;
    col = l.cols[(size_t)lo * l.W + (i - l.row_start[lo])];
}

// Row and column group of a workgroup of a (groups per row) x (rows) grid.
// Workgroups go round-robin over the 8 XCDs by their linear index, so the
// groups of one row are given consecutive slots of one XCD (where the block
// count is a multiple of 8): what they share of the row (candidate
// descriptors, right-view columns) comes into that XCD's L2 once instead of
// once per group.
__device__ __forceinline__ void xcd_row_group(int& row, int& group) {
    const int nb = gridDim.x, total = nb * gridDim.y;
    const int lin = blockIdx.x + nb * blockIdx.y;
    const int q = total % 8 == 0 ? (lin % 8) * (total / 8) + lin / 8 : lin;
    row = q / nb;
    group = q - row * nb;
}

// The same with the rows dealt out over the XCDs (XCD x takes rows x, x + 8,
// ..., each row's groups on consecutive slots of its XCD; where the row count
// is a multiple of 8): the row's shared data still comes into one L2, and
// every XCD gets rows from every part of the image, so work that varies with
// the image region stays balanced over the XCDs.
__device__ __forceinline__ void xcd_row_group_il(int& row, int& group) {
    const int nb = gridDim.x, H = gridDim.y;
    const int lin = blockIdx.x + nb * blockIdx.y;
    if (H % 8 == 0) {
        const int x = lin % 8, j = lin / 8, rb = j / nb;
        row = 8 * rb + x;
        group = j - rb * nb;
    } else {
        row = blockIdx.y;
        group = blockIdx.x;
    }
}

__device__ __forceinline__ uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
    // lowered to v_med3_u32
    return max(min(a, b), min(max(a, b), c));
}

__device__ __forceinline__ int imed3(int a, int b, int c) {
    // lowered to v_med3_i32
    return max(min(a, b), min(max(a, b), c));
}

}  // namespace bicos_hip
