/*
 * BICOS C++ API types for the MI355X (gfx950) build.
 *
 * Same names and meaning as libBICOS v2.2.0 include/common.hpp:32-90, with
 * OpenCV removed: `Image` is this library's own device image (the role
 * cv::cuda::GpuMat plays in the reference's CUDA build, common.hpp:50-56).
 * Precision gains HALF_MFMA (FP16 matrix-core NCC candidates with FP32
 * rescoring) and MOMENTS (closed-form candidates from exact sample moments
 * with FP32 rescoring), which the reference does not have.
 */
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <optional>
#include <string>
#include <type_traits>
#include <variant>

namespace BICOS {

using uint128_t = __uint128_t;

/* include/common.hpp:34-48 */
template<typename T>
constexpr T INVALID_DISP =
    std::numeric_limits<T>::has_quiet_NaN ? std::numeric_limits<T>::quiet_NaN()
                                          : std::numeric_limits<T>::lowest();

template<typename T>
constexpr bool is_invalid(T disparity) {
    if constexpr (std::is_floating_point_v<T>)
        return std::isnan(disparity);
    else
        return disparity == INVALID_DISP<T>;
}

/* element type codes (OpenCV depth codes) */
enum ImageType : int { U8 = 0, U16 = 2, S16 = 3, F32 = 5, F64 = 6 };

/* A device-resident single-channel image: rows x cols elements, rows `step`
 * bytes apart. Copies share the buffer (reference counted), like GpuMat. */
class Image {
public:
    Image() = default;
    Image(int rows, int cols, int type) { create(rows, cols, type); }

    /* (Re)allocates unless the image already has this size and type. */
    void create(int rows, int cols, int type);
    void release();

    /* Synchronous copies between host memory (host_step bytes per row) and
     * the device image; upload() creates the image first. */
    void upload(const void* host, int rows, int cols, int type, size_t host_step = 0);
    void download(void* host, size_t host_step = 0) const;

    bool empty() const { return data == nullptr; }
    int type() const { return type_; }
    int depth() const { return type_; }
    size_t elemSize() const;

    void* data = nullptr;
    int rows = 0, cols = 0;
    size_t step = 0;

private:
    int type_ = -1;
    std::shared_ptr<void> mem_;
};

enum class TransformMode { LIMITED, FULL };
enum class Precision { SINGLE, DOUBLE, HALF_MFMA, MOMENTS };

namespace Variant {
    struct NoDuplicates {};
    struct Consistency {
        int max_lr_diff = 1;
        bool no_dupes = false;
    };
} // namespace Variant

using SearchVariant = std::variant<Variant::NoDuplicates, Variant::Consistency>;

/* include/common.hpp:73-82 */
struct Config {
    std::optional<float> nxcorr_threshold = 0.5f;
    std::optional<float> subpixel_step = std::nullopt;
    std::optional<float> min_variance = std::nullopt;
    TransformMode mode = TransformMode::LIMITED;
    Precision precision = Precision::SINGLE;
    SearchVariant variant = Variant::NoDuplicates {};
};

/* include/common.hpp:84-90 */
class Exception: public std::exception {
    std::string _message;

public:
    Exception(const std::string& message): _message(message) {}
    const char* what() const throw() override { return _message.c_str(); }
};

} // namespace BICOS
