// Point clouds from a disparity map: the reprojection the reference's CLI
// asks OpenCV for (src/cli.cpp:228-250: cv::reprojectImageTo3D(disp, points,
// Q, false, CV_32F)) and the selection its save_pointcloud applies before
// writing (include/fileutils.hpp:45-90), on the device.
//
// Reprojection (OpenCV's published definition, no missing-value handling):
// for the pixel in column x, row y with disparity d (16-bit disparities are
// whole pixels), [X Y Z W] = Q [x y d 1] in double precision and the point is
// (X / W, Y / W, Z / W) stored as float. The sums are formed as OpenCV's loop
// forms them: the row part (Q[i][1] y + Q[i][3]) plus the column part
// Q[i][0] x, plus Q[i][2] d, then multiplied by 1 / W.
//
// Selection: pixels with a valid disparity whose three coordinates are
// finite and (unless negative z is allowed) whose z is not negative, in
// row-major order. Rows are counted, their prefix sum places every row's
// points, and each row is written in column order.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/bicos_hip.h"
#include "host_common.hpp"

namespace bicos_hip {
namespace {

struct QMat {
    double q[16];  // row-major
};

template<typename TD>
__device__ __forceinline__ bool disp_valid(TD d);
template<>
__device__ __forceinline__ bool disp_valid<int16_t>(int16_t d) { return d != INVALID16; }
template<>
__device__ __forceinline__ bool disp_valid<float>(float d) { return !__builtin_isnan(d); }

template<typename TD>
__global__ __launch_bounds__(256) void k_reproject(const TD* __restrict__ disp, int64_t dstep, int W,
                                                   const QMat Q, float* __restrict__ pts) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W)
        return;
    const double d = (double)disp[(size_t)y * dstep + x];
    const double* q = Q.q;
    const double X = (q[1] * y + q[3]) + q[0] * x + q[2] * d;
    const double Yv = (q[5] * y + q[7]) + q[4] * x + q[6] * d;
    const double Z = (q[9] * y + q[11]) + q[8] * x + q[10] * d;
    const double Wv = (q[13] * y + q[15]) + q[12] * x + q[14] * d;
    const double iW = 1.0 / Wv;
    float* p = pts + ((size_t)y * W + x) * 3;
    p[0] = (float)(X * iW);
    p[1] = (float)(Yv * iW);
    p[2] = (float)(Z * iW);
}

// what save_pointcloud does with the point of pixel (x, y): 0 skip (invalid
// disparity), 1 write, 2 skip non-finite, 3 skip negative z
template<typename TD>
__device__ __forceinline__ int pc_class(const TD* disp, int64_t dstep, const float* pts, int W, int x,
                                        int y, bool allow_negative_z) {
    if (!disp_valid<TD>(disp[(size_t)y * dstep + x]))
        return 0;
    const float* p = pts + ((size_t)y * W + x) * 3;
    if (!__builtin_isfinite(p[0]) || !__builtin_isfinite(p[1]) || !__builtin_isfinite(p[2]))
        return 2;
    if (!allow_negative_z && p[2] < 0.0f)
        return 3;
    return 1;
}

// row y = blockIdx.x: points written in the row; skipped totals by atomics
template<typename TD>
__global__ __launch_bounds__(256) void k_pc_count(const TD* __restrict__ disp, int64_t dstep,
                                                  const float* __restrict__ pts, int W, int allow_neg,
                                                  int* __restrict__ row_count,
                                                  unsigned long long* __restrict__ skipped) {
    __shared__ int part[3][4];
    const int y = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int cnt[3] = {0, 0, 0};
    for (int x = threadIdx.x; x < W; x += 256) {
        const int c = pc_class<TD>(disp, dstep, pts, W, x, y, allow_neg != 0);
        if (c) ++cnt[c - 1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int v = cnt[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) part[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) tot[k] = part[k][0] + part[k][1] + part[k][2] + part[k][3];
        row_count[y] = tot[0];
        if (tot[1]) atomicAdd(&skipped[0], (unsigned long long)tot[1]);
        if (tot[2]) atomicAdd(&skipped[1], (unsigned long long)tot[2]);
    }
}

// row y = blockIdx.x: its points from row_start[y] on, in column order
template<typename TD>
__global__ __launch_bounds__(256) void k_pc_write(const TD* __restrict__ disp, int64_t dstep,
                                                  const float* __restrict__ pts, int W, int allow_neg,
                                                  const int* __restrict__ row_start,
                                                  float* __restrict__ xyz) {
    __shared__ int wsum[4];
    const int y = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = row_start[y];
    for (int x0 = 0; x0 < W; x0 += 256) {
        const int x = x0 + threadIdx.x;
        const bool take = x < W && pc_class<TD>(disp, dstep, pts, W, x, y, allow_neg != 0) == 1;
        const uint64_t m = __ballot(take);
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        if (take) {
            int off = base + __popcll(m & ((1ull << lane) - 1));
            for (int w = 0; w < wave; ++w) off += wsum[w];
            const float* p = pts + ((size_t)y * W + x) * 3;
            float* o = xyz + (size_t)off * 3;
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
        }
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0ull;
}

int check_disparity(const bicos_image* d) {
    if (!d || !d->data)
        return fail(BICOS_ERR_EXCEPTION, "disparity: null image");
    if (d->type != BICOS_16S && d->type != BICOS_32F)
        return fail(BICOS_ERR_EXCEPTION, "disparity: must be BICOS_16S or BICOS_32F");
    if (d->rows < 1 || d->cols < 1)
        return fail(BICOS_ERR_EXCEPTION, "disparity: empty image");
    const int es = d->type == BICOS_16S ? 2 : 4;
    if (d->step % es != 0 || d->step < (int64_t)d->cols * es)
        return fail(BICOS_ERR_EXCEPTION, "disparity: bad row step");
    return BICOS_OK;
}

}  // namespace
}  // namespace bicos_hip

using namespace bicos_hip;

extern "C" {

int bicos_hip_reproject(const bicos_image* disparity, const double* Q, float* points, void* stream) {
    int rc;
    if ((rc = check_disparity(disparity))) return rc;
    if (!Q || !points) return fail(BICOS_ERR_INVALID_ARGUMENT, "reproject: null Q or points");
    QMat q;
    for (int i = 0; i < 16; ++i) q.q[i] = Q[i];
    const hipStream_t s = (hipStream_t)stream;
    const int W = disparity->cols, H = disparity->rows;
    const dim3 grid((W + 255) / 256, H), block(256);
    if (disparity->type == BICOS_16S)
        LAUNCH("reproject", k_reproject<int16_t>, grid, block, 0, s, (const int16_t*)disparity->data,
               disparity->step / 2, W, q, points);
    else
        LAUNCH("reproject", k_reproject<float>, grid, block, 0, s, (const float*)disparity->data,
               disparity->step / 4, W, q, points);
    return BICOS_OK;
}

int bicos_hip_pointcloud(const bicos_image* disparity, const float* points, int32_t allow_negative_z,
                         float* xyz, int64_t* counts, void* stream) {
    int rc;
    if ((rc = check_disparity(disparity))) return rc;
    if (!points || !xyz || !counts) return fail(BICOS_ERR_INVALID_ARGUMENT, "pointcloud: null buffer");
    const hipStream_t s = (hipStream_t)stream;
    const int W = disparity->cols, H = disparity->rows;
    int dev;
    if ((rc = current_device(&dev))) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    // row counts, their prefix sums (row_count + H) and the two skipped totals
    const size_t cbytes = list_counts_bytes(H);
    void* ws;
    if ((rc = ws_acquire(dev, cbytes + 256, s, &ws, WS_POINTS))) return rc;
    int* row_count = (int*)ws;
    unsigned long long* skipped = (unsigned long long*)((uint8_t*)ws + cbytes);
    auto body = [&]() -> int {
        LAUNCH("pointcloud", k_zero_u64, dim3(1), dim3(64), 0, s, skipped, 2);
        if (disparity->type == BICOS_16S)
            LAUNCH("pointcloud", k_pc_count<int16_t>, dim3(H), dim3(256), 0, s, (const int16_t*)disparity->data,
                   disparity->step / 2, points, W, (int)allow_negative_z, row_count, skipped);
        else
            LAUNCH("pointcloud", k_pc_count<float>, dim3(H), dim3(256), 0, s, (const float*)disparity->data,
                   disparity->step / 4, points, W, (int)allow_negative_z, row_count, skipped);
        int r;
        if ((r = run_row_scan(row_count, H, s))) return r;
        if (disparity->type == BICOS_16S)
            LAUNCH("pointcloud", k_pc_write<int16_t>, dim3(H), dim3(256), 0, s, (const int16_t*)disparity->data,
                   disparity->step / 2, points, W, (int)allow_negative_z, (const int*)(row_count + H), xyz);
        else
            LAUNCH("pointcloud", k_pc_write<float>, dim3(H), dim3(256), 0, s, (const float*)disparity->data,
                   disparity->step / 4, points, W, (int)allow_negative_z, (const int*)(row_count + H), xyz);
        int written = 0;
        unsigned long long sk[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(&written, row_count + 2 * H, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(sk, skipped, sizeof(sk), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        counts[0] = written;
        counts[1] = (int64_t)sk[0];
        counts[2] = (int64_t)sk[1];
        return BICOS_OK;
    };
    rc = body();
    const int rc2 = ws_release(dev, s, WS_POINTS);
    return rc ? rc : rc2;
}

}  // extern "C"
