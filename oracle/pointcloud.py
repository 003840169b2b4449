"""CPU restatement of the reference CLI's point-cloud path (test
infrastructure only: used by tests/, never by the product).

The reference reprojects with OpenCV (third-party, `dependency('opencv4')`,
unversioned, meson.build:31; absent here):
    cv::reprojectImageTo3D(disp, points, Q, false, CV_32F)     src/cli.cpp:239
whose published definition is, for the pixel in column x and row y with
disparity d (16-bit disparities taken as whole pixels),
    [X Y Z W]^T = Q [x y d 1]^T,   point = (X/W, Y/W, Z/W)  (double, stored as float).
OpenCV's loop forms each sum as (Q[i][1] y + Q[i][3]) + Q[i][0] x + Q[i][2] d
and multiplies by 1/W; `reproject` below does the same in float64.

It then writes the points with save_pointcloud (include/fileutils.hpp:45-90):
pixels with an invalid disparity are skipped (is_invalid: -32768 for int16,
NaN for float; include/common.hpp:34-48), then points with a non-finite
coordinate (counted), then, unless negative z is allowed, points with z < 0
(counted); the rest in row-major order.

Pinned in tests/test_oracle.py by closed forms of a rectified-rig Q
(z = f * baseline / d), an identity-like Q and hand-counted selections.
"""
import numpy as np

INVALID16 = -32768


def reproject(disp: np.ndarray, Q) -> np.ndarray:
    """(rows, cols) int16/float32 disparity -> (rows, cols, 3) float32 points."""
    q = np.asarray(Q, dtype=np.float64).reshape(4, 4)
    rows, cols = disp.shape
    y = np.arange(rows, dtype=np.float64)[:, None]
    x = np.arange(cols, dtype=np.float64)[None, :]
    d = disp.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        comp = [(q[i, 1] * y + q[i, 3]) + q[i, 0] * x + q[i, 2] * d for i in range(4)]
        iw = 1.0 / comp[3]
        out = np.stack([(comp[i] * iw).astype(np.float32) for i in range(3)], axis=-1)
    return out


def select(disp: np.ndarray, points: np.ndarray, allow_negative_z: bool = False):
    """-> ((N, 3) float32 points in row-major order, written, nonfinite, negative_z)."""
    if disp.dtype == np.int16:
        valid = disp != INVALID16
    else:
        valid = ~np.isnan(disp)
    finite = np.isfinite(points).all(axis=-1)
    nonfinite = valid & ~finite
    negz = valid & finite & (points[..., 2] < 0) if not allow_negative_z else np.zeros_like(valid)
    keep = valid & finite & ~negz
    return points[keep], int(keep.sum()), int(nonfinite.sum()), int(negz.sum())
