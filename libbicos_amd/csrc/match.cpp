// C++ API (include/bicos/match.hpp) on top of the C ABI (include/bicos_hip.h).
//
// Mirrors src/lib.cpp:31-49 + the output-allocation rules of
// src/impl/cuda.cu.in:139-226 / src/impl/cpu.cpp:79-91, and turns C error
// codes back into the reference's exception types.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "bicos/match.hpp"
#include "bicos_hip.h"

namespace BICOS {

namespace {

[[noreturn]] void throw_code(int rc) {
    const char* msg = bicos_hip_last_error();
    if (rc == BICOS_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
    throw Exception(msg);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Exception(std::string(what) + ": " + hipGetErrorString(e));
}

size_t elem_size_of(int type) {
    switch (type) {
        case U8: return 1;
        case U16: case S16: return 2;
        case F32: return 4;
        case F64: return 8;
        default: throw Exception("unsupported image type");
    }
}

bicos_image to_c(const Image& img) {
    bicos_image c{};
    c.data = img.data;
    c.rows = img.rows;
    c.cols = img.cols;
    c.type = img.type();
    c.step = (int64_t)img.step;
    return c;
}

} // namespace

size_t Image::elemSize() const { return elem_size_of(type_); }

void Image::create(int r, int c, int t) {
    if (data && rows == r && cols == c && type_ == t)
        return;
    release();
    const size_t es = elem_size_of(t);
    void* p = nullptr;
    size_t pitch = 0;
    hip_check(hipMallocPitch(&p, &pitch, (size_t)c * es, (size_t)r), "hipMallocPitch");
    mem_ = std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
    data = p;
    rows = r;
    cols = c;
    step = pitch;
    type_ = t;
}

void Image::release() {
    mem_.reset();
    data = nullptr;
    rows = cols = 0;
    step = 0;
    type_ = -1;
}

void Image::upload(const void* host, int r, int c, int t, size_t host_step) {
    create(r, c, t);
    const size_t row_bytes = (size_t)c * elemSize();
    hip_check(hipMemcpy2D(data, step, host, host_step ? host_step : row_bytes, row_bytes, (size_t)r,
                          hipMemcpyHostToDevice),
              "hipMemcpy2D");
}

void Image::download(void* host, size_t host_step) const {
    const size_t row_bytes = (size_t)cols * elemSize();
    hip_check(hipMemcpy2D(host, host_step ? host_step : row_bytes, data, step, row_bytes,
                          (size_t)rows, hipMemcpyDeviceToHost),
              "hipMemcpy2D");
}

void match(const std::vector<Image>& stack0, const std::vector<Image>& stack1, Image& disparity,
           Config cfg, Image* corrmap, void* stream) {
    if (stack0.empty() || stack1.empty())
        throw Exception("need at least two images");
    if (stack0.size() != stack1.size())
        throw Exception("both stacks need the same number of images");

    bicos_config c;
    bicos_hip_config_default(&c);
    c.has_nxcorr_threshold = cfg.nxcorr_threshold.has_value();
    c.nxcorr_threshold = cfg.nxcorr_threshold.value_or(0.0f);
    c.has_subpixel_step = cfg.subpixel_step.has_value();
    c.subpixel_step = cfg.subpixel_step.value_or(0.0f);
    c.has_min_variance = cfg.min_variance.has_value();
    c.min_variance = cfg.min_variance.value_or(0.0f);
    c.mode = cfg.mode == TransformMode::FULL ? BICOS_MODE_FULL : BICOS_MODE_LIMITED;
    c.precision = cfg.precision == Precision::DOUBLE      ? BICOS_PRECISION_DOUBLE
                : cfg.precision == Precision::HALF_MFMA ? BICOS_PRECISION_HALF_MFMA
                : cfg.precision == Precision::MOMENTS   ? BICOS_PRECISION_MOMENTS
                                                        : BICOS_PRECISION_SINGLE;
    if (std::holds_alternative<Variant::Consistency>(cfg.variant)) {
        const auto& v = std::get<Variant::Consistency>(cfg.variant);
        c.variant = BICOS_VARIANT_CONSISTENCY;
        c.max_lr_diff = v.max_lr_diff;
        c.no_dupes = v.no_dupes;
    } else {
        c.variant = BICOS_VARIANT_NO_DUPLICATES;
    }

    std::vector<bicos_image> s0, s1;
    for (const auto& im : stack0) s0.push_back(to_c(im));
    for (const auto& im : stack1) s1.push_back(to_c(im));

    const int rows = stack0.front().rows, cols = stack0.front().cols;
    disparity.create(rows, cols, bicos_hip_disparity_type(&c));
    bicos_image d = to_c(disparity);
    bicos_image cm{};
    bicos_image* pcm = nullptr;
    const int ctype = bicos_hip_corrmap_type(&c);
    if (corrmap && ctype >= 0) {
        corrmap->create(rows, cols, ctype);
        cm = to_c(*corrmap);
        pcm = &cm;
    }

    const int rc = bicos_hip_match(s0.data(), s1.data(), (uint32_t)s0.size(), &c, &d, pcm, stream);
    if (rc != BICOS_OK)
        throw_code(rc);
}

} // namespace BICOS
