"""libbicos_amd — BICOS binary correspondence search on MI355X (gfx950).

Python host side over the C ABI of libbicos_hip.so (include/bicos_hip.h).
The module surface mirrors the reference's pybind11 module `pybicos`
(src/pybicos.cpp:102-177): ``Config``, ``TransformMode``, ``Precision``,
``Variant.NoDuplicates``, ``Variant.Consistency``, ``Exception``,
``invalid_disparity`` and ``match(stack0, stack1, cfg) -> (disparity, corrmap)``
on host (numpy) images. ``match_device`` runs on device-resident torch tensors
without host copies; the stage functions (``transform``, ``search``,
``agree``, ``agree_subpixel``) expose the individual kernels.

There is no CPU fallback: importing the package fails if the native library
is missing, and every call runs the gfx950 kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence, Union

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbicos_hip.so")
SCENE_LIB_PATH = os.path.join(HERE, "libbicos_scene.so")

# element type codes (OpenCV depth codes, include/bicos_hip.h)
T_8U, T_16U, T_16S, T_32F, T_64F = 0, 2, 3, 5, 6
OK, ERR_EXCEPTION, ERR_INVALID_ARGUMENT, ERR_HIP = 0, 1, 2, 3


class Exception(RuntimeError):  # noqa: A001 - mirrors pybicos.Exception
    """BICOS::Exception (include/common.hpp:84-90)."""


class _Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("rows", C.c_int32), ("cols", C.c_int32),
                ("type", C.c_int32), ("_pad", C.c_int32), ("step", C.c_int64)]


class _Config(C.Structure):
    _fields_ = [("has_nxcorr_threshold", C.c_int32), ("nxcorr_threshold", C.c_float),
                ("has_subpixel_step", C.c_int32), ("subpixel_step", C.c_float),
                ("has_min_variance", C.c_int32), ("min_variance", C.c_float),
                ("mode", C.c_int32), ("precision", C.c_int32), ("variant", C.c_int32),
                ("max_lr_diff", C.c_int32), ("no_dupes", C.c_int32)]


class _WindowOptions(C.Structure):
    _fields_ = [("in_scale", C.c_float), ("p_conf", C.c_float), ("hw_lo", C.c_int32),
                ("hw_hi", C.c_int32), ("bins", C.c_int32), ("fallback", C.c_int32)]


class _WindowStats(C.Structure):
    _fields_ = [("pixels", C.c_int64), ("fallback_pixels", C.c_int64)]


class _KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double)]


def _load():
    # PyTorch-ROCm bundles its own HIP runtime under the same soname
    # (libamdhip64.so.7). Whichever copy a process loads first is the one
    # both use; loading /opt/rocm's copy first leaves torch's device
    # initialisation and ours on different runtimes, and the device is not
    # found. So torch, when present, is imported before the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m libbicos_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, f32 = C.c_void_p, C.c_int32, C.c_uint32, C.c_float
    P = C.POINTER
    L.bicos_hip_last_error.restype = C.c_char_p
    L.bicos_hip_config_default.argtypes = [P(_Config)]
    L.bicos_hip_disparity_type.argtypes = [P(_Config)]
    L.bicos_hip_corrmap_type.argtypes = [P(_Config)]
    L.bicos_hip_descriptor_bits.argtypes = [u32, i32]
    L.bicos_hip_match.argtypes = [P(_Image), P(_Image), u32, P(_Config), P(_Image), P(_Image), vp]
    L.bicos_hip_match_noalloc.argtypes = [P(_Image), P(_Image), u32, P(_Config), P(_Image), P(_Image), vp]
    L.bicos_hip_noalloc_scratch_bytes.restype = C.c_size_t
    L.bicos_hip_transform.argtypes = [P(_Image), u32, i32, i32, vp, vp]
    L.bicos_hip_search.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, vp]
    L.bicos_hip_agree.argtypes = [vp, P(_Image), P(_Image), u32, f32, i32, f32, i32, vp, vp]
    L.bicos_hip_agree_subpixel.argtypes = [vp, P(_Image), P(_Image), u32, f32, f32, i32, f32,
                                           i32, vp, vp, vp]
    L.bicos_hip_set_search_impl.argtypes = [i32]
    L.bicos_hip_search_chunk_cols.argtypes = [i32, i32]
    L.bicos_hip_set_subpixel_impl.argtypes = [i32]
    L.bicos_hip_timing_enable.argtypes = [i32]
    L.bicos_hip_timing_read.argtypes = [P(_KernelStat), i32]
    L.bicos_hip_malloc.argtypes = [C.c_size_t]
    L.bicos_hip_malloc.restype = vp
    L.bicos_hip_free.argtypes = [vp]
    L.bicos_hip_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
    L.bicos_hip_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
    L.bicos_hip_window_options_default.argtypes = [P(_WindowOptions)]
    L.bicos_hip_window_options_default.restype = None
    L.bicos_hip_window_init.argtypes = [vp, C.c_uint64]
    L.bicos_hip_set_cnn_impl.argtypes = [i32]
    L.bicos_hip_window_label.argtypes = [P(_Image), i32, vp, vp]
    L.bicos_hip_window_loss_grad.argtypes = [P(_Image), P(_Image), u32, vp, P(_WindowOptions), vp,
                                             vp, vp, vp]
    L.bicos_hip_window_adam.argtypes = [vp, vp, vp, vp, vp, f32, f32, f32, f32, i32, vp]
    L.bicos_hip_window_predict.argtypes = [P(_Image), P(_Image), u32, vp, P(_WindowOptions), vp,
                                           vp, vp, vp]
    L.bicos_hip_search_windowed.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp]
    L.bicos_hip_match_windowed.argtypes = [P(_Image), P(_Image), u32, P(_Config), vp,
                                           P(_WindowOptions), P(_Image), P(_Image),
                                           P(_WindowStats), vp]
    d = C.c_double
    L.bicos_hip_scene_generate.argtypes = [i32, i32, i32, i32, d, d, C.c_uint64, d, d, d, d, d,
                                           vp, vp, vp, vp]
    L.bicos_hip_reproject.argtypes = [P(_Image), P(C.c_double), vp, vp]
    L.bicos_hip_pointcloud.argtypes = [P(_Image), vp, i32, vp, P(C.c_int64), vp]
    return L


_lib = _load()

EXPORTED_SYMBOLS = (
    "bicos_hip_config_default", "bicos_hip_disparity_type", "bicos_hip_corrmap_type",
    "bicos_hip_descriptor_bits", "bicos_hip_match", "bicos_hip_match_noalloc",
    "bicos_hip_noalloc_scratch_bytes", "bicos_hip_last_error",
    "bicos_hip_release_workspace", "bicos_hip_set_search_impl", "bicos_hip_search_chunk_cols",
    "bicos_hip_set_subpixel_impl",
    "bicos_hip_transform", "bicos_hip_search", "bicos_hip_agree",
    "bicos_hip_agree_subpixel", "bicos_hip_timing_enable", "bicos_hip_timing_read",
    "bicos_hip_malloc", "bicos_hip_free", "bicos_hip_memcpy_h2d", "bicos_hip_memcpy_d2h",
    "bicos_hip_device_synchronize", "bicos_hip_window_param_count",
    "bicos_hip_window_options_default", "bicos_hip_window_init", "bicos_hip_window_label",
    "bicos_hip_window_loss_grad", "bicos_hip_window_adam", "bicos_hip_window_predict",
    "bicos_hip_match_windowed", "bicos_hip_search_windowed", "bicos_hip_scene_generate", "bicos_hip_set_cnn_impl",
    "bicos_hip_reproject", "bicos_hip_pointcloud",
)


def lib():
    return _lib


def _check(rc: int) -> None:
    if rc == OK:
        return
    msg = _lib.bicos_hip_last_error().decode()
    if rc == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)  # std::invalid_argument in the reference
    raise Exception(msg)


# ---------------------------------------------------------------- API types

class TransformMode:
    """include/common.hpp:58 (pybicos: TransformMode.Full / .Limited)."""
    Limited = 0
    Full = 1
    LIMITED = 0
    FULL = 1


class Precision:
    """include/common.hpp:59-61; HalfMFMA, Moments and MomentsFast have no
    reference counterpart (subpixel candidates scored from FP16 matrix-core
    products, winners rescored in FP32; resp. in closed form from exact
    moments, with the reference's accept/reject decisions (Moments) or
    decided on the closed form's own values (MomentsFast))."""
    Single = 0
    Double = 1
    HalfMFMA = 2
    Moments = 3
    MomentsFast = 4


class Variant:
    """include/common.hpp:63-71."""

    @dataclass(frozen=True)
    class NoDuplicates:
        def __repr__(self):
            return "NoDuplicates"

    @dataclass(frozen=True)
    class Consistency:
        max_lr_diff: int = 1
        no_dupes: bool = False


@dataclass
class Config:
    """BICOS::Config with the reference defaults (include/common.hpp:73-82)."""
    nxcorr_threshold: Optional[float] = 0.5
    subpixel_step: Optional[float] = None
    min_variance: Optional[float] = None
    mode: int = TransformMode.Limited
    precision: int = Precision.Single
    variant: Union[Variant.NoDuplicates, Variant.Consistency] = field(
        default_factory=Variant.NoDuplicates)

    def _c(self) -> _Config:
        c = _Config()
        _lib.bicos_hip_config_default(C.byref(c))
        c.has_nxcorr_threshold = int(self.nxcorr_threshold is not None)
        c.nxcorr_threshold = float(self.nxcorr_threshold or 0.0)
        c.has_subpixel_step = int(self.subpixel_step is not None)
        c.subpixel_step = float(self.subpixel_step or 0.0)
        c.has_min_variance = int(self.min_variance is not None)
        c.min_variance = float(self.min_variance or 0.0)
        c.mode = int(self.mode)
        c.precision = int(self.precision)
        if isinstance(self.variant, Variant.Consistency):
            c.variant = 1
            c.max_lr_diff = int(self.variant.max_lr_diff)
            c.no_dupes = int(bool(self.variant.no_dupes))
        elif isinstance(self.variant, Variant.NoDuplicates):
            c.variant = 0
        else:
            raise ValueError("unimplemented")
        return c


def invalid_disparity(dtype) -> Union[int, float]:
    """INVALID_DISP<T> (include/common.hpp:34-37): -32768 for int16, NaN for float32."""
    dt = np.dtype(dtype)
    if dt == np.int16:
        return -32768
    if dt == np.float32:
        return float("nan")
    raise Exception(f"invalid_disparity undefined for {dt}")


def descriptor_bits(n: int, mode: int) -> int:
    return _lib.bicos_hip_descriptor_bits(n, mode)


def disparity_dtype(cfg: Config):
    return np.float32 if _lib.bicos_hip_disparity_type(C.byref(cfg._c())) == T_32F else np.int16


def corrmap_dtype(cfg: Config):
    t = _lib.bicos_hip_corrmap_type(C.byref(cfg._c()))
    return None if t < 0 else (np.float64 if t == T_64F else np.float32)


_NP2T = {np.dtype(np.uint8): T_8U, np.dtype(np.uint16): T_16U, np.dtype(np.int16): T_16S,
         np.dtype(np.float32): T_32F, np.dtype(np.float64): T_64F}


# --------------------------------------------------------- device (torch)

def _torch():
    import torch
    return torch


def _tdtype_code(t) -> int:
    torch = _torch()
    m = {torch.uint8: T_8U, torch.uint16: T_16U, torch.int16: T_16S, torch.float32: T_32F,
         torch.float64: T_64F}
    if t.dtype not in m:
        raise Exception(f"unsupported tensor dtype {t.dtype}")
    return m[t.dtype]


def _img_of_tensor2d(t) -> _Image:
    if t.dim() != 2 or t.stride(1) != 1:
        raise Exception("images must be 2-D with unit column stride")
    return _Image(t.data_ptr(), t.shape[0], t.shape[1], _tdtype_code(t), 0,
                  t.stride(0) * t.element_size())


def _stack_images(stack) -> "C.Array":
    """stack: (n, rows, cols) tensor or a sequence of (rows, cols) tensors.
    The array keeps the tensors alive (`_keep`) and records their device."""
    if hasattr(stack, "dim"):
        if stack.dim() != 3 or stack.stride(2) != 1:
            raise Exception("images must be 2-D with unit column stride")
        # frame descriptors straight from the base pointer and strides (no
        # per-frame tensor views: this runs on every call)
        n, rows, cols = stack.shape
        es = stack.element_size()
        base, fstride, step, code = stack.data_ptr(), stack.stride(0) * es, stack.stride(1) * es, _tdtype_code(stack)
        arr = (_Image * n)(*[_Image(base + i * fstride, rows, cols, code, 0, step) for i in range(n)])
        arr._keep = (stack,)
        arr._device = stack.device
        return arr
    frames = list(stack)
    arr = (_Image * len(frames))(*[_img_of_tensor2d(f) for f in frames])
    arr._keep = frames
    arr._device = frames[0].device if frames else None
    return arr


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def match_device(stack0, stack1, cfg: Optional[Config] = None, disparity=None, corrmap=None,
                 want_corrmap: bool = False, stream=None, noalloc: bool = False):
    """BICOS::match on device tensors. stack0/stack1: (n, rows, cols) uint8/uint16
    CUDA tensors (or sequences of 2-D tensors). Returns (disparity, corrmap|None)
    as CUDA tensors; preallocated outputs of the right dtype are reused.
    noalloc: the library allocates nothing for the call
    (bicos_hip_match_noalloc: its static per-GPU scratch; calls that overlap
    on one GPU race)."""
    torch = _torch()
    cfg = cfg or Config()
    c = cfg._c()
    s0, s1 = _stack_images(stack0), _stack_images(stack1)
    n = len(s0)
    if len(s1) != n:
        raise Exception("both stacks need the same number of images")
    rows, cols = s0[0].rows, s0[0].cols
    dev = s0._device
    dt = torch.float32 if _lib.bicos_hip_disparity_type(C.byref(c)) == T_32F else torch.int16
    if disparity is None or disparity.dtype != dt or tuple(disparity.shape) != (rows, cols):
        disparity = torch.empty((rows, cols), dtype=dt, device=dev)
    ct = _lib.bicos_hip_corrmap_type(C.byref(c))
    cimg = None
    if (want_corrmap or corrmap is not None) and ct >= 0:
        cdt = torch.float64 if ct == T_64F else torch.float32
        if corrmap is None or corrmap.dtype != cdt or tuple(corrmap.shape) != (rows, cols):
            corrmap = torch.empty((rows, cols), dtype=cdt, device=dev)
        cimg = _img_of_tensor2d(corrmap)
    else:
        corrmap = None
    dimg = _img_of_tensor2d(disparity)
    entry = _lib.bicos_hip_match_noalloc if noalloc else _lib.bicos_hip_match
    _check(entry(s0, s1, n, C.byref(c), C.byref(dimg), C.byref(cimg) if cimg is not None else None,
                 _stream_ptr(stream)))
    return disparity, corrmap


def reproject(disparity, Q, stream=None):
    """cv::reprojectImageTo3D(disparity, points, Q, false, CV_32F) as the
    reference CLI calls it (src/cli.cpp:239): (rows, cols, 3) float32 points
    of an int16 or float32 CUDA disparity tensor; Q: 4x4 (row-major)."""
    torch = _torch()
    q = np.ascontiguousarray(np.asarray(Q, dtype=np.float64).reshape(16))
    img = _img_of_tensor2d(disparity)
    pts = torch.empty((disparity.shape[0], disparity.shape[1], 3), dtype=torch.float32,
                      device=disparity.device)
    _check(_lib.bicos_hip_reproject(C.byref(img), q.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.c_void_p(pts.data_ptr()), _stream_ptr(stream)))
    return pts


def pointcloud(disparity, Q, allow_negative_z: bool = False, stream=None):
    """The points save_pointcloud writes (include/fileutils.hpp:45-90): those of
    valid disparities with finite coordinates (and z >= 0 unless allowed), in
    row-major order. Returns ((N, 3) float32 CUDA tensor, {"written",
    "nonfinite", "negative_z"})."""
    torch = _torch()
    pts = reproject(disparity, Q, stream)
    img = _img_of_tensor2d(disparity)
    xyz = torch.empty((disparity.numel(), 3), dtype=torch.float32, device=disparity.device)
    counts = (C.c_int64 * 3)()
    _check(_lib.bicos_hip_pointcloud(C.byref(img), C.c_void_p(pts.data_ptr()), int(bool(allow_negative_z)),
                                     C.c_void_p(xyz.data_ptr()), counts, _stream_ptr(stream)))
    n = int(counts[0])
    return xyz[:n], {"written": n, "nonfinite": int(counts[1]), "negative_z": int(counts[2])}


def transform(stack, mode: int, words: Optional[int] = None, out=None, stream=None):
    """Descriptor transform of one view -> (rows, cols, words) int32 tensor (uint32 bits)."""
    torch = _torch()
    imgs = _stack_images(stack)
    n = len(imgs)
    rows, cols = imgs[0].rows, imgs[0].cols
    if words is None:
        bits = descriptor_bits(n, mode)
        if bits < 0:
            raise ValueError("input stacks too large")
        words = bits // 32
    if out is None:
        out = torch.empty((rows, cols, words), dtype=torch.int32, device=imgs._device)
    _check(_lib.bicos_hip_transform(imgs, n, mode, words, C.c_void_p(out.data_ptr()),
                                    _stream_ptr(stream)))
    return out


def search(desc0, desc1, flags: int, max_lr_diff: int = -1, out=None, stream=None):
    """Row search on (rows, cols, words) descriptor tensors -> int16 disparity."""
    torch = _torch()
    rows, cols, words = desc0.shape
    if out is None:
        out = torch.empty((rows, cols), dtype=torch.int16, device=desc0.device)
    _check(_lib.bicos_hip_search(C.c_void_p(desc0.data_ptr()), C.c_void_p(desc1.data_ptr()), rows,
                                 cols, words, flags, max_lr_diff, C.c_void_p(out.data_ptr()),
                                 _stream_ptr(stream)))
    return out


def agree(disparity, stack0, stack1, threshold: float, min_variance: Optional[float] = None,
          precision: int = Precision.Single, corrmap=None, stream=None):
    """In-place NCC agreement on an int16 disparity tensor (min_variance per frame)."""
    s0, s1 = _stack_images(stack0), _stack_images(stack1)
    _check(_lib.bicos_hip_agree(C.c_void_p(disparity.data_ptr()), s0, s1, len(s0),
                                float(threshold), int(min_variance is not None),
                                float(min_variance or 0.0), int(precision),
                                C.c_void_p(corrmap.data_ptr()) if corrmap is not None else None,
                                _stream_ptr(stream)))
    return disparity


def agree_subpixel(disparity, stack0, stack1, threshold: float, step: float,
                   min_variance: Optional[float] = None, precision: int = Precision.Single,
                   out=None, corrmap=None, stream=None):
    torch = _torch()
    s0, s1 = _stack_images(stack0), _stack_images(stack1)
    if out is None:
        out = torch.empty(disparity.shape, dtype=torch.float32, device=disparity.device)
    _check(_lib.bicos_hip_agree_subpixel(
        C.c_void_p(disparity.data_ptr()), s0, s1, len(s0), float(threshold), float(step),
        int(min_variance is not None), float(min_variance or 0.0), int(precision),
        C.c_void_p(out.data_ptr()),
        C.c_void_p(corrmap.data_ptr()) if corrmap is not None else None, _stream_ptr(stream)))
    return out


def timing_enable(on: bool = True) -> None:
    _check(_lib.bicos_hip_timing_enable(int(on)))


def timing_read() -> dict:
    buf = (_KernelStat * 64)()
    k = _lib.bicos_hip_timing_read(buf, 64)
    if k < 0:
        _check(k)
    return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(min(k, 64))}


SEARCH_MFMA, SEARCH_VALU, SEARCH_MFMA_I8 = 0, 1, 2


def set_search_impl(impl: int) -> None:
    """Row-search kernel for later calls: SEARCH_MFMA (default: fp4 matrix-core
    products for descriptors of 64 bits and more on rows up to 4096 columns,
    int8 otherwise), SEARCH_MFMA_I8 (int8 products throughout) or SEARCH_VALU."""
    _check(_lib.bicos_hip_set_search_impl(int(impl)))


def search_chunk_cols(impl: int, words: int) -> int:
    """Candidate columns per LDS chunk of the full-row search that `impl`
    runs for `words`-word descriptors (0 for SEARCH_VALU)."""
    k = _lib.bicos_hip_search_chunk_cols(int(impl), int(words))
    if k < 0:
        raise ValueError("unknown search implementation or descriptor width")
    return k


SUBPIXEL_PRUNED, SUBPIXEL_SCAN = 0, 1


def set_subpixel_impl(impl: int) -> None:
    """Subpixel kernel at Precision.SINGLE / DOUBLE for later calls:
    SUBPIXEL_PRUNED (default; closed-form bounds pick the grid points the
    exact arithmetic scores) or SUBPIXEL_SCAN (every grid point). Identical
    results."""
    _check(_lib.bicos_hip_set_subpixel_impl(int(impl)))


def release_workspace() -> None:
    _check(_lib.bicos_hip_release_workspace())


# ------------------------------------------------------------ host (numpy)

class _DevBuf:
    def __init__(self, nbytes: int):
        self.p = _lib.bicos_hip_malloc(max(int(nbytes), 1))
        if not self.p:
            raise Exception(_lib.bicos_hip_last_error().decode())

    def __del__(self):
        if getattr(self, "p", None):
            _lib.bicos_hip_free(self.p)
            self.p = None


def match(stack0: Sequence[np.ndarray], stack1: Sequence[np.ndarray], cfg: Optional[Config] = None):
    """pybicos.match (src/pybicos.cpp:102-119): lists of 2-D uint8/uint16 numpy
    images in, (disparity, corrmap) numpy arrays out; corrmap is None when no
    threshold is set. Uploads, runs on the GPU, downloads."""
    cfg = cfg or Config()
    if len(stack0) == 0 or len(stack1) == 0:
        raise Exception("empty stacks")
    c = cfg._c()
    bufs, imgs0, imgs1 = [], [], []
    for src, dst in ((stack0, imgs0), (stack1, imgs1)):
        for a in src:
            a = np.ascontiguousarray(a)
            if a.ndim != 2 or a.dtype not in (np.uint8, np.uint16):
                raise Exception(f"unimplemented cvtype_of(dtype={a.dtype})")
            b = _DevBuf(a.nbytes)
            _check(_lib.bicos_hip_memcpy_h2d(b.p, a.ctypes.data, a.nbytes))
            bufs.append(b)
            dst.append(_Image(b.p, a.shape[0], a.shape[1], _NP2T[a.dtype], 0, a.strides[0]))
    rows, cols = imgs0[0].rows, imgs0[0].cols
    ddt = np.float32 if _lib.bicos_hip_disparity_type(C.byref(c)) == T_32F else np.int16
    disp = np.empty((rows, cols), ddt)
    dbuf = _DevBuf(disp.nbytes)
    dimg = _Image(dbuf.p, rows, cols, _NP2T[np.dtype(ddt)], 0, cols * disp.itemsize)
    ct = _lib.bicos_hip_corrmap_type(C.byref(c))
    corr, cbuf, cimg = None, None, None
    if ct >= 0:
        corr = np.empty((rows, cols), np.float64 if ct == T_64F else np.float32)
        cbuf = _DevBuf(corr.nbytes)
        cimg = _Image(cbuf.p, rows, cols, ct, 0, cols * corr.itemsize)
    a0 = (_Image * len(imgs0))(*imgs0)
    a1 = (_Image * len(imgs1))(*imgs1)
    if len(imgs1) < len(imgs0):
        raise Exception("both stacks need the same number of images")
    _check(_lib.bicos_hip_match(a0, a1, len(imgs0), C.byref(c), C.byref(dimg),
                                C.byref(cimg) if cimg is not None else None, None))
    _check(_lib.bicos_hip_memcpy_d2h(disp.ctypes.data, dbuf.p, disp.nbytes))
    if corr is not None:
        _check(_lib.bicos_hip_memcpy_d2h(corr.ctypes.data, cbuf.p, corr.nbytes))
    return disp, corr
