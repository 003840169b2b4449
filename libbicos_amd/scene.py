"""Synthetic rectified multi-shot stereo scenes (libbicos_scene.so).

The reference benchmarks on a downloaded dataset (data/prepare.sh:20) that is
unavailable offline; these scenes stand in for it. See csrc/scene.cpp for the
construction (speckle frames, smooth disparity field with planar steps,
Gaussian noise, a Bernoulli set of constant dark pixels).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libbicos_scene.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m libbicos_amd.build`")
        L = C.CDLL(path)
        L.bicos_scene_generate.argtypes = ([C.c_int] * 4 + [C.c_double] * 2 + [C.c_uint64]
                                           + [C.c_double] * 5 + [C.c_int] + [C.c_void_p] * 3)
        _lib = L
    return _lib


@dataclass(frozen=True)
class SceneSpec:
    width: int
    height: int
    frames: int
    depth: int          # 8 or 16 bits per sample
    dmin: float
    dmax: float
    lo: float           # speckle value range
    hi: float
    sigma: float        # sensor noise
    p_dark: float = 0.05
    dark_value: float = 10.0


# The workloads named by the benchmark configurations.
SCENE_640x480_U8 = SceneSpec(640, 480, 12, 8, 8.0, 72.0, 20.0, 235.0, 2.0, 0.05, 10.0)
SCENE_1920x1200_U16 = SceneSpec(1920, 1200, 33, 16, 0.0, 256.0, 1000.0, 60000.0, 150.0, 0.05, 500.0)
SCENE_4096x3000_U16 = SceneSpec(4096, 3000, 64, 16, 0.0, 512.0, 1000.0, 60000.0, 150.0, 0.05, 500.0)

BENCH_SEED = 0x600DF00D  # bench/cuda.cu:39 of the reference


def generate(spec: SceneSpec, seed: int, threads: int = 0, height: int | None = None,
             width: int | None = None, frames: int | None = None):
    """Returns (left, right, disparity): (n, H, W) uint8/uint16 stacks and the
    (H, W) float32 ground-truth disparity of the left view."""
    W = width or spec.width
    H = height or spec.height
    n = frames or spec.frames
    dt = np.uint8 if spec.depth == 8 else np.uint16
    left = np.empty((n, H, W), dt)
    right = np.empty((n, H, W), dt)
    disp = np.empty((H, W), np.float32)
    rc = _load().bicos_scene_generate(W, H, n, spec.depth // 8, spec.dmin, spec.dmax,
                                      seed & 0xFFFFFFFFFFFFFFFF, spec.lo, spec.hi, spec.sigma,
                                      spec.p_dark, spec.dark_value, int(threads),
                                      left.ctypes.data, right.ctypes.data, disp.ctypes.data)
    if rc != 0:
        raise ValueError("bad scene parameters")
    return left, right, disp
