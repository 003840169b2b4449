/*
 * BICOS::match for the MI355X (gfx950) build.
 *
 * Reference: libBICOS v2.2.0 include/match.hpp:31-41 (CUDA build variant):
 * same argument order and meaning; the CUDA build's cv::cuda::Stream& becomes
 * a hipStream_t passed as void* (nullptr = default stream).
 *
 * `disparity` is (re)created as S16, or F32 when both a threshold and a
 * subpixel step are set; `corrmap`, when given and a threshold is set, is
 * (re)created as F32 (F64 for Precision::DOUBLE) and NaN where no
 * correlation was evaluated. Work is enqueued on `stream`.
 * Errors: BICOS::Exception, or std::invalid_argument for too-deep stacks,
 * unknown variants and non-positive subpixel steps (src/impl/cpu.cpp:151).
 */
#pragma once

#include "common.hpp"

#include <vector>

namespace BICOS {

void match(
    const std::vector<Image>& stack0,
    const std::vector<Image>& stack1,
    Image& disparity,
    Config cfg = Config {},
    Image* corrmap = nullptr,
    void* stream = nullptr
);

} // namespace BICOS
