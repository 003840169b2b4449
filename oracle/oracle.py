"""ctypes front-end of the BICOS CPU oracle (bicos_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product package never
imports it.

Stacks are numpy arrays of shape (n, rows, cols), dtype uint8 or uint16
(frame-planar, like the list of n single-channel images BICOS::match takes).
Descriptors are (rows, cols, words) uint32, bit k in word k // 32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class OracleConfig(C.Structure):
    _fields_ = [
        ("has_thr", C.c_int), ("thr", C.c_float),
        ("has_step", C.c_int), ("step", C.c_float),
        ("has_minvar", C.c_int), ("minvar", C.c_float),
        ("full", C.c_int), ("dbl", C.c_int),
        ("consistency", C.c_int), ("max_lr_diff", C.c_int), ("no_dupes", C.c_int),
    ]


def build(force: bool = False) -> str:
    args = ["make", "-s", "-C", HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_transform.argtypes = [vp, i64, i32, i32, i32, i32, i32, i32, vp]
        L.oracle_search.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp]
        L.oracle_nxcorr_u16.argtypes = [vp, vp, i32, i32, f32]
        L.oracle_nxcorr_u16.restype = C.c_float
        L.oracle_nxcorr_u8.argtypes = [vp, vp, i32, i32, f32]
        L.oracle_nxcorr_u8.restype = C.c_float
        L.oracle_nxcorrd_u16.argtypes = [vp, vp, i32, i32, f32]
        L.oracle_nxcorrd_u16.restype = C.c_double
        L.oracle_agree.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, f32, i32, f32, i32, vp]
        L.oracle_agree_subpixel.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, f32, f32, i32,
                                            f32, i32, vp, vp]
        L.oracle_subpixel_grid.argtypes = [f32, vp, i32]
        L.oracle_match.argtypes = [vp, vp, i64, i32, i32, i32, i32, C.POINTER(OracleConfig), vp,
                                   vp, C.POINTER(C.c_int), vp]
        L.oracle_search_window.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, vp, vp]
        L.oracle_match_windowed.argtypes = [vp, vp, i64, i32, i32, i32, i32, C.POINTER(OracleConfig),
                                            vp, vp, i32, i32, vp, vp, C.POINTER(C.c_int), vp,
                                            C.POINTER(C.c_int64)]
        L.oracle_interp_sample.argtypes = [i32, i32, i32, i32, f32]
        L.oracle_subpixel_curve.argtypes = [i32, i32, i64, vp, vp, f32, i32, vp]
        L.oracle_subpixel_closed_form.argtypes = [i32, i32, i64, vp, vp, f32, i32, vp, vp]
        L.oracle_required_bits.argtypes = [i32, i32]
        L.oracle_descriptor_words.argtypes = [i32, i32]
        L.oracle_set_threads.argtypes = [i32]
        # the process's CPU share, at most 16 threads (a GPU box's share per GPU)
        L.oracle_set_threads(min(16, len(os.sched_getaffinity(0))))
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def _frames(stack: np.ndarray):
    stack = np.ascontiguousarray(stack)
    if stack.ndim != 3 or stack.dtype not in (np.uint8, np.uint16):
        raise ValueError("stack must be (n, rows, cols) uint8/uint16")
    n = stack.shape[0]
    base = stack.ctypes.data
    fstride = stack.strides[0]
    ptrs = (C.c_void_p * n)(*[base + t * fstride for t in range(n)])
    return stack, ptrs, stack.strides[1], stack.dtype.itemsize


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


def required_bits(n: int, full: bool) -> int:
    return lib().oracle_required_bits(n, int(full))


def descriptor_words(n: int, full: bool) -> int:
    return lib().oracle_descriptor_words(n, int(full))


def transform(stack: np.ndarray, full: bool, words: int | None = None) -> np.ndarray:
    stack, ptrs, step, depth = _frames(stack)
    n, rows, cols = stack.shape
    if words is None:
        words = descriptor_words(n, full)
    out = np.zeros((rows, cols, words), np.uint32)
    _check(lib().oracle_transform(ptrs, step, depth, n, rows, cols, int(full), words,
                                  out.ctypes.data))
    return out


def search(d0: np.ndarray, d1: np.ndarray, flags: int, max_lr_diff: int = -1) -> np.ndarray:
    d0 = np.ascontiguousarray(d0, np.uint32)
    d1 = np.ascontiguousarray(d1, np.uint32)
    rows, cols, words = d0.shape
    out = np.empty((rows, cols), np.int16)
    _check(lib().oracle_search(d0.ctypes.data, d1.ctypes.data, rows, cols, words, flags,
                               max_lr_diff, out.ctypes.data))
    return out


def nxcorr(p0: np.ndarray, p1: np.ndarray, minvar: float | None = None) -> float:
    """FP32 NCC of two sample vectors; minvar already scaled by n."""
    p0 = np.ascontiguousarray(p0)
    p1 = np.ascontiguousarray(p1, p0.dtype)
    f = lib().oracle_nxcorr_u16 if p0.dtype == np.uint16 else lib().oracle_nxcorr_u8
    return f(p0.ctypes.data, p1.ctypes.data, len(p0), int(minvar is not None),
             float(minvar or 0.0))


def nxcorrd(p0: np.ndarray, p1: np.ndarray, minvar: float | None = None) -> float:
    p0 = np.ascontiguousarray(p0, np.uint16)
    p1 = np.ascontiguousarray(p1, np.uint16)
    return lib().oracle_nxcorrd_u16(p0.ctypes.data, p1.ctypes.data, len(p0),
                                    int(minvar is not None), float(minvar or 0.0))


def agree(disp: np.ndarray, s0: np.ndarray, s1: np.ndarray, thr: float,
          minvar: float | None = None, double: bool = False, corrmap: bool = False):
    """In-place-style agree: returns (new int16 disparity, corrmap or None).
    minvar is already multiplied by n."""
    s0, p0, step, depth = _frames(s0)
    s1, p1, _, _ = _frames(s1)
    n, rows, cols = s0.shape
    d = np.array(disp, np.int16, copy=True, order="C")
    cm = None
    if corrmap:
        cm = np.full((rows, cols), np.nan, np.float64 if double else np.float32)
    _check(lib().oracle_agree(d.ctypes.data, p0, p1, step, depth, n, rows, cols, float(thr),
                              int(minvar is not None), float(minvar or 0.0), int(double),
                              cm.ctypes.data if cm is not None else None))
    return d, cm


def agree_subpixel(disp: np.ndarray, s0: np.ndarray, s1: np.ndarray, thr: float, step_x: float,
                   minvar: float | None = None, double: bool = False, corrmap: bool = False):
    s0, p0, step, depth = _frames(s0)
    s1, p1, _, _ = _frames(s1)
    n, rows, cols = s0.shape
    d = np.ascontiguousarray(disp, np.int16)
    out = np.empty((rows, cols), np.float32)
    cm = None
    if corrmap:
        cm = np.full((rows, cols), np.nan, np.float64 if double else np.float32)
    _check(lib().oracle_agree_subpixel(d.ctypes.data, p0, p1, step, depth, n, rows, cols,
                                       float(thr), float(step_x), int(minvar is not None),
                                       float(minvar or 0.0), int(double), out.ctypes.data,
                                       cm.ctypes.data if cm is not None else None))
    return out, cm


def subpixel_grid(step: float) -> np.ndarray:
    buf = np.zeros(4096, np.float32)
    k = lib().oracle_subpixel_grid(float(step), buf.ctypes.data, len(buf))
    if k < 0:
        raise ValueError("bad step")
    return buf[:k].copy()


def interp_sample(y0: int, y1: int, y2: int, x: float, dtype=np.uint16) -> int:
    """One interpolated subpixel sample (agree.hpp:152-164), cast to dtype."""
    return lib().oracle_interp_sample(1 if np.dtype(dtype) == np.uint8 else 2, int(y0), int(y1),
                                      int(y2), float(x))


def subpixel_curve(left: np.ndarray, right3: np.ndarray, step: float) -> np.ndarray:
    """FP32 subpixel correlation of every grid point (agree.hpp:143-171, no
    min_variance) for a batch of pixels: left (P, n), right3 (P, 3, n) of
    uint8/uint16 -> (P, npts) float32."""
    left = np.ascontiguousarray(left)
    right3 = np.ascontiguousarray(right3, left.dtype)
    P, n = left.shape
    assert right3.shape == (P, 3, n)
    npts = len(subpixel_grid(step))
    out = np.empty((P, npts), np.float32)
    _check(lib().oracle_subpixel_curve(left.dtype.itemsize, n, P, left.ctypes.data,
                                       right3.ctypes.data, float(step), npts, out.ctypes.data))
    return out


def subpixel_closed_form(left: np.ndarray, right3: np.ndarray, step: float):
    """The closed-form correlation of MOMENTS_FAST (unrounded parabola
    samples, double precision) at every grid point for a batch of pixels:
    left (P, n), right3 (P, 3, n) -> (nxc (P, npts), var1 (P, npts))."""
    left = np.ascontiguousarray(left)
    right3 = np.ascontiguousarray(right3, left.dtype)
    P, n = left.shape
    npts = len(subpixel_grid(step))
    out = np.empty((P, npts), np.float64)
    var1 = np.empty((P, npts), np.float64)
    _check(lib().oracle_subpixel_closed_form(left.dtype.itemsize, n, P, left.ctypes.data,
                                             right3.ctypes.data, float(step), npts, out.ctypes.data,
                                             var1.ctypes.data))
    return out, var1


def match(s0: np.ndarray, s1: np.ndarray, *, thr: float | None = 0.5, step: float | None = None,
          min_variance: float | None = None, full: bool = False, double: bool = False,
          consistency: tuple[int, bool] | None = None, corrmap: bool = False):
    """BICOS::match on the CPU (cpu.cpp:96-155). Returns (disparity, corrmap|None):
    int16 unless thr and step are both set, then float32."""
    s0, p0, rstep, depth = _frames(s0)
    s1, p1, _, _ = _frames(s1)
    n, rows, cols = s0.shape
    cfg = OracleConfig(
        int(thr is not None), float(thr or 0.0),
        int(step is not None), float(step or 0.0),
        int(min_variance is not None), float(min_variance or 0.0),
        int(full), int(double),
        int(consistency is not None),
        int(consistency[0]) if consistency else 1,
        int(consistency[1]) if consistency else 0,
    )
    d16 = np.empty((rows, cols), np.int16)
    df = np.empty((rows, cols), np.float32)
    cm = None
    if corrmap and thr is not None:
        cm = np.empty((rows, cols), np.float64 if double else np.float32)
    is_float = C.c_int(0)
    _check(lib().oracle_match(p0, p1, rstep, depth, n, rows, cols, C.byref(cfg), d16.ctypes.data,
                              df.ctypes.data, C.byref(is_float),
                              cm.ctypes.data if cm is not None else None))
    return (df if is_float.value else d16), cm


def _cells(center: np.ndarray, halfwidth: np.ndarray, rows: int, cols: int):
    center = np.ascontiguousarray(center, np.float32)
    halfwidth = np.ascontiguousarray(halfwidth, np.int16)
    if center.shape != (rows // 4, cols // 4) or halfwidth.shape != center.shape:
        raise ValueError("center/halfwidth must be (rows // 4, cols // 4) cell arrays")
    return center, halfwidth


def search_window(d0: np.ndarray, d1: np.ndarray, flags: int, max_lr_diff: int,
                  center: np.ndarray, halfwidth: np.ndarray) -> np.ndarray:
    """Row search limited to each 4 x 4 cell's window [col - rint(c) - hw,
    col - rint(c) + hw] (bicos.hpp:50-113 restricted; the Consistency back
    search over [col - hw, col + hw])."""
    d0 = np.ascontiguousarray(d0, np.uint32)
    d1 = np.ascontiguousarray(d1, np.uint32)
    rows, cols, words = d0.shape
    center, halfwidth = _cells(center, halfwidth, rows, cols)
    out = np.empty((rows, cols), np.int16)
    _check(lib().oracle_search_window(d0.ctypes.data, d1.ctypes.data, rows, cols, words, flags,
                                      max_lr_diff, center.ctypes.data, halfwidth.ctypes.data,
                                      out.ctypes.data))
    return out


def match_windowed(s0: np.ndarray, s1: np.ndarray, center: np.ndarray, halfwidth: np.ndarray, *,
                   hw_lo: int = 6, fallback: int = 1, thr: float | None = 0.5,
                   step: float | None = None, min_variance: float | None = None,
                   full: bool = False, double: bool = False,
                   consistency: tuple[int, bool] | None = None, corrmap: bool = False):
    """The learned-window BICOS::match on the CPU for given cell windows:
    windowed search, NCC stage, then (fallback 1 = all, 2 = unsure) the
    full-row search and NCC stage of the fallback set. Returns (disparity,
    corrmap | None, fallback pixel count)."""
    s0, p0, rstep, depth = _frames(s0)
    s1, p1, _, _ = _frames(s1)
    n, rows, cols = s0.shape
    center, halfwidth = _cells(center, halfwidth, rows, cols)
    cfg = OracleConfig(
        int(thr is not None), float(thr or 0.0),
        int(step is not None), float(step or 0.0),
        int(min_variance is not None), float(min_variance or 0.0),
        int(full), int(double),
        int(consistency is not None),
        int(consistency[0]) if consistency else 1,
        int(consistency[1]) if consistency else 0,
    )
    d16 = np.empty((rows, cols), np.int16)
    df = np.empty((rows, cols), np.float32)
    cm = None
    if corrmap and thr is not None:
        cm = np.empty((rows, cols), np.float64 if double else np.float32)
    is_float = C.c_int(0)
    nfb = C.c_int64(0)
    _check(lib().oracle_match_windowed(p0, p1, rstep, depth, n, rows, cols, C.byref(cfg),
                                       center.ctypes.data, halfwidth.ctypes.data, int(hw_lo),
                                       int(fallback), d16.ctypes.data, df.ctypes.data,
                                       C.byref(is_float), cm.ctypes.data if cm is not None else None,
                                       C.byref(nfb)))
    return (df if is_float.value else d16), cm, int(nfb.value)
