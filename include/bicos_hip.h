/*
 * bicos_hip.h — C ABI of the MI355X (gfx950) BICOS matching library.
 *
 * This is the drop-in boundary for libBICOS's matching path. Every entry
 * point takes plain pointers and sizes; device buffers are HIP device
 * pointers, streams are hipStream_t passed as void*. Nothing here depends on
 * OpenCV, torch or any C++ type. Calls are asynchronous on the given stream
 * (NULL = the default stream): inputs must stay alive and outputs must not be
 * read until the stream has reached the end of the call.
 *
 * Reference interfaces replaced (libBICOS v2.2.0):
 *   bicos_hip_match            BICOS::match            include/match.hpp:31-41,
 *                                                      src/lib.cpp:31-49,
 *                                                      src/impl/cuda.cu.in:233-292
 *   bicos_config               BICOS::Config           include/common.hpp:58-82
 *   bicos_hip_transform        transform_{limited,full}_kernel
 *                                                      include/impl/cuda/descriptor_transform.cuh:30-149
 *   bicos_hip_search           bicos_kernel(_smem)     include/impl/cuda/bicos.cuh:77-176
 *   bicos_hip_agree            agree_kernel            include/impl/cuda/agree.cuh:110-159
 *   bicos_hip_agree_subpixel   agree_subpixel_kernel   include/impl/cuda/agree.cuh:161-259
 *   error codes                BICOS::Exception / std::invalid_argument
 *                                                      include/common.hpp:84-90, src/impl/cpu.cpp:151
 */
#ifndef BICOS_HIP_H
#define BICOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element type codes: OpenCV's depth codes, kept so callers can pass
 * cv::Mat::depth() straight through (include/common.hpp:50-56 uses cv::Mat). */
enum {
    BICOS_8U = 0,
    BICOS_16U = 2,
    BICOS_16S = 3,
    BICOS_32F = 5,
    BICOS_64F = 6
};

/* Return codes. BICOS_ERR_EXCEPTION corresponds to BICOS::Exception,
 * BICOS_ERR_INVALID_ARGUMENT to std::invalid_argument in the reference. */
enum {
    BICOS_OK = 0,
    BICOS_ERR_EXCEPTION = 1,
    BICOS_ERR_INVALID_ARGUMENT = 2,
    BICOS_ERR_HIP = 3
};

/* TransformMode (include/common.hpp:58) */
enum { BICOS_MODE_LIMITED = 0, BICOS_MODE_FULL = 1 };

/* Precision (include/common.hpp:59-61). HALF_MFMA, MOMENTS and
 * MOMENTS_FAST have no reference counterpart and concern the subpixel stage
 * (the single-point agreement stays the exact FP32 one); none is bit-exact
 * (see the precision tests):
 *   HALF_MFMA     grid points scored from FP16 matrix-core cross products
 *                 and FP64 means/variances, winners rescored in FP32;
 *   MOMENTS       grid points scored in closed form from exact (FP64) sample
 *                 moments; accept/reject is the reference's (an error bound
 *                 sends undecided pixels to the exact FP32 stage);
 *   MOMENTS_FAST  the same closed form from exact moments, decided on its
 *                 own FP32 values (no exact stage): pixels within about 1e-4
 *                 of the threshold can be decided the other way. Needs at most 64
 *                 grid points and no corrmap (else the exact FP32 stage runs). */
enum {
    BICOS_PRECISION_SINGLE = 0,
    BICOS_PRECISION_DOUBLE = 1,
    BICOS_PRECISION_HALF_MFMA = 2,
    BICOS_PRECISION_MOMENTS = 3,
    BICOS_PRECISION_MOMENTS_FAST = 4
};

/* SearchVariant (include/common.hpp:63-71) */
enum { BICOS_VARIANT_NO_DUPLICATES = 0, BICOS_VARIANT_CONSISTENCY = 1 };

/* One single-channel image, device-resident (or host, where stated).
 * Like cv::cuda::GpuMat: rows x cols elements of `type`, rows `step` bytes apart. */
typedef struct bicos_image {
    void* data;
    int32_t rows;
    int32_t cols;
    int32_t type;   /* BICOS_8U, BICOS_16U, BICOS_16S, BICOS_32F, BICOS_64F */
    int32_t _pad;
    int64_t step;   /* bytes between rows */
} bicos_image;

/* BICOS::Config with explicit has_* flags for each std::optional.
 * Reference defaults (include/common.hpp:73-82): has_nxcorr_threshold = 1,
 * nxcorr_threshold = 0.5, no subpixel, no min_variance, LIMITED, SINGLE,
 * NoDuplicates; Consistency defaults max_lr_diff = 1, no_dupes = 0. */
typedef struct bicos_config {
    int32_t has_nxcorr_threshold;
    float nxcorr_threshold;
    int32_t has_subpixel_step;
    float subpixel_step;
    int32_t has_min_variance;
    float min_variance;
    int32_t mode;        /* BICOS_MODE_* */
    int32_t precision;   /* BICOS_PRECISION_* */
    int32_t variant;     /* BICOS_VARIANT_* */
    int32_t max_lr_diff; /* Consistency only */
    int32_t no_dupes;    /* Consistency only */
} bicos_config;

/* Fills *cfg with the reference defaults. */
void bicos_hip_config_default(bicos_config* cfg);

/* Element type match() writes to `disparity` for this config: BICOS_16S, or
 * BICOS_32F when both a threshold and a subpixel step are set
 * (src/impl/cpu.cpp:85-88). */
int bicos_hip_disparity_type(const bicos_config* cfg);

/* Element type of the correlation map for this config (BICOS_32F, or
 * BICOS_64F for BICOS_PRECISION_DOUBLE), or -1 when no threshold is set and
 * no map is produced (src/impl/cuda.cu.in:183-191). */
int bicos_hip_corrmap_type(const bicos_config* cfg);

/* Descriptor width in bits chosen for n frames (32/64/128/256), or -1 when
 * too many bits would be needed (src/impl/cpu.cpp:118-152). */
int bicos_hip_descriptor_bits(uint32_t n, int32_t mode);

/* BICOS::match. stack0/stack1: n device images each (BICOS_8U or BICOS_16U,
 * all the same size and type). disparity: device image of the caller,
 * rows x cols of bicos_hip_disparity_type(cfg). corrmap: NULL or a device
 * image of bicos_hip_corrmap_type(cfg); it is filled with NaN and then
 * written for every pixel whose correlation is evaluated (CPU semantics,
 * src/impl/cpu.cpp:80-83). stream: hipStream_t or NULL. The library keeps a
 * per-device workspace between calls (grown on demand, released by
 * bicos_hip_release_workspace); calls on different streams are ordered
 * through it. Returns BICOS_OK or an error code (bicos_hip_last_error()). */
int bicos_hip_match(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                    const bicos_config* cfg, bicos_image* disparity, bicos_image* corrmap,
                    void* stream);

/* bicos_hip_match without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * semantics and results as bicos_hip_match; the caller still passes the
 * stacks and outputs. The call's scratch (descriptor planes, the int16
 * disparity before the subpixel stage, reverse-search columns, the MOMENTS
 * status, packed frames for stacks that are not one uniformly spaced block)
 * comes from a static per-GPU array of bicos_hip_noalloc_scratch_bytes()
 * bytes (96 MiB; 1920x1200 with 128-bit descriptors needs about 80 MB),
 * held by every GPU the library runs on until the process ends; a call
 * needing more is rejected with BICOS_ERR_INVALID_ARGUMENT before anything
 * is launched. All calls on one GPU share that array: calls that overlap on
 * one GPU (on different streams, or from several threads) race and must be
 * ordered by the caller. Calls no hipMalloc/hipFree; the first call on a
 * device looks up the array's address there once (hipGetSymbolAddress).
 * Per-kernel timing (bicos_hip_timing_enable) creates events when it is on. */
int bicos_hip_match_noalloc(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                            const bicos_config* cfg, bicos_image* disparity, bicos_image* corrmap,
                            void* stream);

/* Size in bytes of the per-GPU scratch array of bicos_hip_match_noalloc. */
size_t bicos_hip_noalloc_scratch_bytes(void);

/* Thread-local message of the last failed call on this thread. */
const char* bicos_hip_last_error(void);

/* Frees the cached workspace of the current device (waits for its last use). */
int bicos_hip_release_workspace(void);

/* Selects the row-search kernel for subsequent calls of this process:
 * BICOS_SEARCH_MFMA (default; Hamming costs from matrix-core products: fp4
 * operands for descriptors of 64 bits and more on rows up to 4096 columns,
 * int8 otherwise), BICOS_SEARCH_MFMA_I8 (int8 matrix-core products
 * throughout) or BICOS_SEARCH_VALU (xor/popcount on the vector ALU). All give
 * identical results; the switch exists for testing and measurement. No
 * reference counterpart (the reference picks between its shared-memory and
 * global-memory search kernels by row size, src/impl/cuda.cu.in:149-177). */
enum { BICOS_SEARCH_MFMA = 0, BICOS_SEARCH_VALU = 1, BICOS_SEARCH_MFMA_I8 = 2 };
int bicos_hip_set_search_impl(int32_t impl);

/* Candidate columns per LDS chunk of the full-row search kernel that `impl`
 * runs for descriptors of `words` u32 words on rows up to 4096 columns (the
 * chunk and tile edges are where a row search changes shape; tests probe
 * widths around them). 0 for BICOS_SEARCH_VALU, -1 for unknown arguments.
 * No reference counterpart. */
int bicos_hip_search_chunk_cols(int32_t impl, int32_t words);

/* Selects the kernel of the subpixel stage at Precision::SINGLE and
 * Precision::DOUBLE for subsequent calls of this process:
 * BICOS_SUBPIXEL_PRUNED (default; exact moments and a rigorous closed-form
 * bound pick the grid points that can be the maximum, the reference's FP32
 * or FP64 arithmetic scores only those) or BICOS_SUBPIXEL_SCAN (the
 * reference's arithmetic at every grid point). Both give identical results.
 * PRUNED applies without a corrmap and with at most 64 grid points; SCAN is
 * used otherwise. No reference counterpart (include/impl/cpu/agree.hpp:94-185
 * scans every grid point). */
enum { BICOS_SUBPIXEL_PRUNED = 0, BICOS_SUBPIXEL_SCAN = 1 };
int bicos_hip_set_subpixel_impl(int32_t impl);

/* ---- stage-level entry points (the reference's kernel-level tests drive
 *      its kernels directly: test/descriptor_transform.cu, test/bicos.cu,
 *      test/agree.cu) ---- */

/* Descriptor transform of one view: frames = n device images (8U/16U),
 * out = device buffer of rows*cols*words uint32 (bit k of a pixel's
 * descriptor in word k/32, bit k%32), words = bits/32 for the width chosen
 * by bicos_hip_descriptor_bits(n, mode) unless a wider one is asked for. */
int bicos_hip_transform(const bicos_image* frames, uint32_t n, int32_t mode, int32_t words,
                        uint32_t* out, void* stream);

/* Row search on descriptor buffers (rows*cols*words uint32 each).
 * flags: 1 = no duplicates, 2 = consistency (include/impl/common.hpp:46-47),
 * 3 = both. out: device int16 rows*cols disparity (-32768 = invalid). */
int bicos_hip_search(const uint32_t* desc0, const uint32_t* desc1, int32_t rows, int32_t cols,
                     int32_t words, int32_t flags, int32_t max_lr_diff, int16_t* out,
                     void* stream);

/* In-place NCC agreement on an int16 disparity (device, rows*cols dense).
 * min_variance is the per-frame value (it is multiplied by n here, like
 * src/impl/cpu.cpp:123). corrmap: NULL or device rows*cols of float
 * (precision SINGLE) or double (DOUBLE), written where evaluated. */
int bicos_hip_agree(int16_t* disparity, const bicos_image* stack0, const bicos_image* stack1,
                    uint32_t n, float nxcorr_threshold, int32_t has_min_variance,
                    float min_variance, int32_t precision, void* corrmap, void* stream);

/* Subpixel agreement: int16 disparity in, float disparity out (NaN =
 * invalid); arguments as bicos_hip_agree. */
int bicos_hip_agree_subpixel(const int16_t* disparity, const bicos_image* stack0,
                             const bicos_image* stack1, uint32_t n, float nxcorr_threshold,
                             float subpixel_step, int32_t has_min_variance, float min_variance,
                             int32_t precision, float* out, void* corrmap, void* stream);

/* ---- learned search window (no reference counterpart) ----
 *
 * A siamese CNN on a 1/4-resolution grid of 4 x 4 cells (input: the 4x4 box
 * means of the first 4 frames of each view minus their mean over those
 * frames, times in_scale; 4 conv 3x3 layers of 32 channels, ReLU between,
 * BF16 weights, FP32 accumulation) correlates left and right features over
 * `bins` horizontal offsets and predicts per cell the centre and half-width
 * of the disparity window the Hamming search then scans. Parameters: a flat
 * device float array of bicos_hip_window_param_count() values, layout
 * w1 b1 w2 b2 w3 b3 w4 b4 with w[out][in][3][3]. */

typedef struct bicos_window_options {
    float in_scale;        /* input scale; 0 = 8 / (largest sample value) */
    float p_conf;          /* peak softmax probability for the narrow window */
    int32_t hw_lo;         /* window half-width (px) of confident cells */
    int32_t hw_hi;         /* ... and of the others */
    int32_t bins;          /* correlation offsets: disparities 4 * (-8 .. bins - 9) px */
    int32_t fallback;      /* BICOS_FALLBACK_*: which rejected pixels get the full-row search */
} bicos_window_options;

/* Full-row search of the pixels whose windowed match the NCC stage rejects
 * (those whose left variance already rejects every match excepted):
 *   NONE    none;
 *   ALL     all of them: a pixel the windowed path rejects is then one the
 *           plain path rejects too;
 *   UNSURE  those in cells the network was unsure about (peak probability
 *           below p_conf: the cells given the hw_hi window). */
enum { BICOS_FALLBACK_NONE = 0, BICOS_FALLBACK_ALL = 1, BICOS_FALLBACK_UNSURE = 2 };

typedef struct bicos_window_stats {
    int64_t pixels;            /* pixels matched */
    int64_t fallback_pixels;   /* pixels sent to the full-row search */
} bicos_window_stats;

int bicos_hip_window_param_count(void);
void bicos_hip_window_options_default(bicos_window_options* opts);

/* Network implementation of predict/match_windowed for subsequent calls:
 * BICOS_CNN_MFMA (default; cell-major BF16 tensors, the four layers in one
 * kernel and the correlation on v_mfma_f32_32x32x16_bf16, bins <= 160),
 * BICOS_CNN_MFMA_LAYERS (the same products, one kernel per layer; equal
 * results bit for bit) or BICOS_CNN_VALU (the training kernels). All compute
 * the same network up to FP32 summation order. */
enum { BICOS_CNN_MFMA = 0, BICOS_CNN_VALU = 1, BICOS_CNN_MFMA_LAYERS = 2 };
int bicos_hip_set_cnn_impl(int32_t impl);

/* Deterministic He-normal initial parameters (biases 0), host memory. */
int bicos_hip_window_init(float* host_params, uint64_t seed);

/* Training labels from a disparity map of the plain path (device image,
 * BICOS_32F with NaN or BICOS_16S with -32768 = invalid): per cell the mean
 * of its valid disparities when at least min_valid of its 16 pixels are
 * valid, else NaN. label: device float, (rows/4) x (cols/4). */
int bicos_hip_window_label(const bicos_image* disparity, int32_t min_valid, float* label,
                           void* stream);

/* Forward pass, softmax cross-entropy against the labels (a label splits
 * linearly over its two neighbouring bins) and backward pass for one stack
 * pair: ADDS the gradient of the summed loss to grad (device, param_count)
 * and the loss sum and labelled-cell count to loss[0], loss[1] (device). */
int bicos_hip_window_loss_grad(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                               const float* params, const bicos_window_options* opts,
                               const float* label, float* grad, float* loss, void* stream);

/* One Adam step on params with moments m, v (device, param_count each):
 * the gradient used is grad / max(count[0], 1) (count: device, e.g. loss + 1);
 * step counts from 1. */
int bicos_hip_window_adam(float* params, float* m, float* v, const float* grad,
                          const float* count, float lr, float beta1, float beta2, float eps,
                          int32_t step, void* stream);

/* Window of every cell: centre (disparity, px) and half-width (px); peak
 * (nullable) receives the peak softmax probability. Device arrays of
 * (rows/4) x (cols/4). */
int bicos_hip_window_predict(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                             const float* params, const bicos_window_options* opts, float* center,
                             int16_t* halfwidth, float* peak, void* stream);

/* BICOS::match with the Hamming search limited to the predicted windows:
 * same arguments, config semantics and outputs as bicos_hip_match, plus the
 * parameters and options. Pixels whose windowed match the NCC stage rejects
 * (and whose left variance does not reject every match) are searched over the
 * full row and their NCC stage is run again (exact FP32 for SINGLE and
 * HALF_MFMA, FP64 for DOUBLE, MOMENTS for MOMENTS). stats: NULL, or host memory filled after the
 * call completes (the call then waits for the stream). Images need at least
 * 4 rows and columns and fewer than 65536 rows. */
int bicos_hip_match_windowed(const bicos_image* stack0, const bicos_image* stack1, uint32_t n,
                             const bicos_config* cfg, const float* params,
                             const bicos_window_options* opts, bicos_image* disparity,
                             bicos_image* corrmap, bicos_window_stats* stats, void* stream);

/* The windowed Hamming search of bicos_hip_match_windowed as a stage (the
 * learned path's counterpart of bicos_hip_search; no reference counterpart,
 * the semantics are include/impl/cpu/bicos.hpp:50-113 restricted to a
 * window). Pixel (row, col) lies in cell (min(row/4, rows/4 - 1),
 * min(col/4, cols/4 - 1)) with centre c and half-width hw (device arrays of
 * (rows/4) x (cols/4)); it searches the candidate columns
 * [k - hw, k + hw] clipped to the row, k = col - rint(c) (half to even), an
 * empty window giving -32768. flags: 1 = no duplicates (within the window),
 * 2 = consistency (the matched column searches the left row back over
 * [col - hw, col + hw]), 3 = both. Descriptor buffers and out as
 * bicos_hip_search. */
int bicos_hip_search_windowed(const uint32_t* desc0, const uint32_t* desc1, int32_t rows,
                              int32_t cols, int32_t words, int32_t flags, int32_t max_lr_diff,
                              const float* center, const int16_t* halfwidth, int16_t* out,
                              void* stream);

/* ---- synthetic scenes (training and test inputs; the host generator's
 *      construction evaluated on the device) ----
 * depth 1 -> uint8, 2 -> uint16 samples; left/right: device n*H*W samples,
 * frame-planar; disp: device H*W float ground truth (nullable). */
int bicos_hip_scene_generate(int32_t width, int32_t height, int32_t n, int32_t depth, double dmin,
                             double dmax, uint64_t seed, double lo, double hi, double sigma,
                             double p_dark, double dark_value, void* left, void* right,
                             float* disp, void* stream);

/* ---- point clouds (the reference CLI's --q-store path, src/cli.cpp:228-250) ---- */

/* cv::reprojectImageTo3D(disparity, points, Q, false, CV_32F) as
 * src/cli.cpp:239 calls it: for the pixel in column x, row y with disparity d,
 * [X Y Z W] = Q [x y d 1] in double precision, point = (X/W, Y/W, Z/W) as
 * float. disparity: device BICOS_16S (whole pixels) or BICOS_32F image; Q:
 * host, 16 doubles row-major; points: device rows*cols*3 floats. */
int bicos_hip_reproject(const bicos_image* disparity, const double* Q, float* points, void* stream);

/* The selection of save_pointcloud (include/fileutils.hpp:45-90): points of
 * valid disparities with finite coordinates and, unless allow_negative_z,
 * z >= 0, packed in row-major order into xyz (device, room for rows*cols*3
 * floats). counts (host, 3 values): points written, skipped as non-finite,
 * skipped for negative z. Waits for the stream (the counts are returned). */
int bicos_hip_pointcloud(const bicos_image* disparity, const float* points, int32_t allow_negative_z,
                         float* xyz, int64_t* counts, void* stream);

/* ---- measurement ---- */

/* Per-kernel GPU time accounting with HIP events recorded around every
 * kernel the library launches (on the stream it launches on). */
typedef struct bicos_kernel_stat {
    char name[48];
    int64_t launches;
    double total_ms;
} bicos_kernel_stat;

int bicos_hip_timing_enable(int32_t enable);   /* resets the statistics */
/* Waits for all recorded events, then copies up to cap entries; returns the
 * number of distinct kernels (or a negative error code). */
int bicos_hip_timing_read(bicos_kernel_stat* out, int32_t cap);

/* ---- device memory helpers for callers without a HIP toolchain ---- */
void* bicos_hip_malloc(size_t bytes);
int bicos_hip_free(void* p);
int bicos_hip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int bicos_hip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int bicos_hip_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* BICOS_HIP_H */
