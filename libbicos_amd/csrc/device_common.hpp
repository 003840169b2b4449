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

constexpr int16_t INVALID16 = -32768;  // INVALID_DISP<int16_t> (include/common.hpp:34-37)

template<typename T>
__device__ __forceinline__ uint32_t load_sample(const FrameTable& f, int t, int row, int col) {
    const T* p = (const T*)(f.ptr[t] + (uint64_t)row * (uint64_t)f.step[t]);
    return (uint32_t)__builtin_nontemporal_load(p + col);
}

template<typename T>
__device__ __forceinline__ uint32_t load_sample_cached(const FrameTable& f, int t, int row, int col) {
    const T* p = (const T*)(f.ptr[t] + (uint64_t)row * (uint64_t)f.step[t]);
    return (uint32_t)p[col];
}

__device__ __forceinline__ uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
    // lowered to v_med3_u32
    return max(min(a, b), min(max(a, b), c));
}

}  // namespace bicos_hip
