"""Learned search window: training and windowed matching.

A siamese CNN (4 conv 3x3 layers, 32 channels, BF16 weights, FP32
accumulation; csrc/cnn.hpp) looks at the first 4 frames of both views on a
1/4-resolution grid, correlates left and right features over horizontal
offsets and predicts for every 4 x 4 cell the centre and half-width of the
disparity window. ``match_windowed_device`` runs BICOS::match with the
Hamming search limited to those windows and sends every pixel whose windowed
match the NCC stage rejects back to the full-row search. The network is
trained on the GPU by the library's own forward/backward kernels
(``bicos_hip_window_loss_grad``, ``bicos_hip_window_adam``) on synthetic
scenes labelled by the plain matching path. No reference counterpart.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass, asdict
from typing import Callable, Optional

import numpy as np

from . import (Config, Exception, T_32F, T_64F, _WindowOptions, _WindowStats, _check,
               _img_of_tensor2d, _lib, _stack_images, _stream_ptr, _torch, match_device)
from . import scene as _scene

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM_COUNT = _lib.bicos_hip_window_param_count()
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "window_1920x1200x33_u16.npz")
CELL = 4          # pixels per cell side
CNN_FRAMES = 4    # frames the network sees


@dataclass
class WindowOptions:
    """bicos_window_options (include/bicos_hip.h)."""
    in_scale: float = 0.0      # 0: 8 / largest sample value
    p_conf: float = 0.5        # peak probability of the narrow window
    hw_lo: int = 6             # half-width (px) of confident cells
    hw_hi: int = 12            # ... and of the others
    bins: int = 80             # correlation offsets: disparities 4 * (-8 .. bins - 9) px
    fallback: int = 1          # FALLBACK_NONE / FALLBACK_ALL / FALLBACK_UNSURE (bools: ALL / NONE)

    def _c(self) -> _WindowOptions:
        fb = self.fallback
        fb = FALLBACK_ALL if fb is True else (FALLBACK_NONE if fb is False else int(fb))
        return _WindowOptions(float(self.in_scale), float(self.p_conf), int(self.hw_lo),
                              int(self.hw_hi), int(self.bins), fb)


# which rejected pixels get the full-row search (include/bicos_hip.h)
FALLBACK_NONE, FALLBACK_ALL, FALLBACK_UNSURE = 0, 1, 2


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() if t is not None else None)


def _check_params(params):
    if params.dtype != _torch().float32 or params.numel() != PARAM_COUNT or not params.is_contiguous():
        raise Exception(f"window parameters must be a contiguous float32 tensor of {PARAM_COUNT}")


CNN_MFMA, CNN_VALU, CNN_MFMA_LAYERS = 0, 1, 2


def set_cnn_impl(impl: int) -> None:
    """Network implementation of predict / match_windowed_device for later
    calls: CNN_MFMA (default, matrix cores, layers fused), CNN_MFMA_LAYERS
    (matrix cores, one kernel per layer) or CNN_VALU (training kernels)."""
    _check(_lib.bicos_hip_set_cnn_impl(int(impl)))


# ----------------------------------------------------------------- weights

def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (bicos_hip_window_init)."""
    p = np.empty(PARAM_COUNT, np.float32)
    _check(_lib.bicos_hip_window_init(p.ctypes.data, seed & 0xFFFFFFFFFFFFFFFF))
    return p


def save(path: str, params: np.ndarray, meta: Optional[dict] = None) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    meta = dict(meta or {})
    np.savez(path, params=np.asarray(params, np.float32),
             meta_keys=np.array(list(meta.keys()), dtype=str),
             meta_values=np.array([repr(v) for v in meta.values()], dtype=str))


def load(path: str = DEFAULT_WEIGHTS):
    """-> (params float32 array, meta dict of strings)"""
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        meta = dict(zip(z["meta_keys"].tolist(), z["meta_values"].tolist()))
    if params.size != PARAM_COUNT:
        raise Exception(f"{path}: {params.size} parameters, expected {PARAM_COUNT}")
    return params, meta


def params_device(params, device="cuda"):
    torch = _torch()
    if isinstance(params, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(device)
    return params.to(device=device, dtype=torch.float32).contiguous()


# ------------------------------------------------------------ entry points

def predict(stack0, stack1, params, opts: Optional[WindowOptions] = None, stream=None):
    """-> (centre float32, half-width int16, peak probability float32), each
    (rows // 4, cols // 4) on the stacks' device."""
    torch = _torch()
    opts = opts or WindowOptions()
    _check_params(params)
    s0, s1 = _stack_images(stack0), _stack_images(stack1)
    n = len(s0)
    Hq, Wq = s0[0].rows // CELL, s0[0].cols // CELL
    dev = s0._device
    center = torch.empty((Hq, Wq), dtype=torch.float32, device=dev)
    hw = torch.empty((Hq, Wq), dtype=torch.int16, device=dev)
    peak = torch.empty((Hq, Wq), dtype=torch.float32, device=dev)
    o = opts._c()
    _check(_lib.bicos_hip_window_predict(s0, s1, n, _ptr(params), C.byref(o), _ptr(center),
                                         _ptr(hw), _ptr(peak), _stream_ptr(stream)))
    return center, hw, peak


def search_windowed(desc0, desc1, flags: int, max_lr_diff: int, center, halfwidth, out=None,
                    stream=None):
    """The windowed row search alone: (rows, cols, words) descriptor tensors
    and (rows // 4, cols // 4) cell windows -> int16 disparity."""
    torch = _torch()
    rows, cols, words = desc0.shape
    if out is None:
        out = torch.empty((rows, cols), dtype=torch.int16, device=desc0.device)
    if tuple(center.shape) != (rows // CELL, cols // CELL) or tuple(halfwidth.shape) != tuple(center.shape):
        raise ValueError("center/halfwidth must be (rows // 4, cols // 4)")
    if center.dtype != torch.float32 or halfwidth.dtype != torch.int16:
        raise ValueError("center must be float32 and halfwidth int16")
    _check(_lib.bicos_hip_search_windowed(_ptr(desc0), _ptr(desc1), rows, cols, words, int(flags),
                                          int(max_lr_diff), _ptr(center.contiguous()),
                                          _ptr(halfwidth.contiguous()), _ptr(out), _stream_ptr(stream)))
    return out


def label(disparity, min_valid: int = 8, stream=None):
    """Cell labels (rows // 4, cols // 4) from a plain-path disparity map."""
    torch = _torch()
    rows, cols = disparity.shape
    out = torch.empty((rows // CELL, cols // CELL), dtype=torch.float32, device=disparity.device)
    img = _img_of_tensor2d(disparity)
    _check(_lib.bicos_hip_window_label(C.byref(img), int(min_valid), _ptr(out), _stream_ptr(stream)))
    return out


def loss_grad(stack0, stack1, params, labels, grad, loss, opts: Optional[WindowOptions] = None,
              stream=None):
    """Adds the summed-loss gradient to grad and (loss sum, count) to loss[0:2]."""
    opts = opts or WindowOptions()
    _check_params(params)
    s0, s1 = _stack_images(stack0), _stack_images(stack1)
    o = opts._c()
    _check(_lib.bicos_hip_window_loss_grad(s0, s1, len(s0), _ptr(params), C.byref(o), _ptr(labels),
                                           _ptr(grad), _ptr(loss), _stream_ptr(stream)))


def adam(params, m, v, grad, count, lr: float, step: int, beta1: float = 0.9,
         beta2: float = 0.999, eps: float = 1e-8, stream=None):
    _check(_lib.bicos_hip_window_adam(_ptr(params), _ptr(m), _ptr(v), _ptr(grad), _ptr(count),
                                      lr, beta1, beta2, eps, int(step), _stream_ptr(stream)))


def match_windowed_device(stack0, stack1, cfg: Optional[Config], params,
                          opts: Optional[WindowOptions] = None, disparity=None, corrmap=None,
                          want_corrmap: bool = False, stats: bool = False, stream=None):
    """BICOS::match with the learned window (bicos_hip_match_windowed).
    Returns (disparity, corrmap|None, stats|None); stats = {"pixels",
    "fallback_pixels"} makes the call wait for the stream."""
    torch = _torch()
    cfg = cfg or Config()
    opts = opts or WindowOptions()
    _check_params(params)
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
    o = opts._c()
    st = _WindowStats()
    _check(_lib.bicos_hip_match_windowed(s0, s1, n, C.byref(c), _ptr(params), C.byref(o),
                                         C.byref(dimg), C.byref(cimg) if cimg is not None else None,
                                         C.byref(st) if stats else None, _stream_ptr(stream)))
    return disparity, corrmap, ({"pixels": st.pixels, "fallback_pixels": st.fallback_pixels}
                                if stats else None)


def scene_device(spec: "_scene.SceneSpec", seed: int, device="cuda", frames: Optional[int] = None,
                 height: Optional[int] = None, width: Optional[int] = None, stream=None):
    """Synthetic scene generated on the device (bicos_hip_scene_generate; the
    host generator's construction) -> (left, right, disparity) tensors.
    Frame t does not depend on the number of frames generated."""
    torch = _torch()
    W, H = width or spec.width, height or spec.height
    n = frames or spec.frames
    dt = torch.uint8 if spec.depth == 8 else torch.uint16
    left = torch.empty((n, H, W), dtype=dt, device=device)
    right = torch.empty((n, H, W), dtype=dt, device=device)
    disp = torch.empty((H, W), dtype=torch.float32, device=device)
    _check(_lib.bicos_hip_scene_generate(W, H, n, spec.depth // 8, spec.dmin, spec.dmax,
                                         seed & 0xFFFFFFFFFFFFFFFF, spec.lo, spec.hi, spec.sigma,
                                         spec.p_dark, spec.dark_value, _ptr(left), _ptr(right),
                                         _ptr(disp), _stream_ptr(stream)))
    return left, right, disp


# ---------------------------------------------------------------- training

# Plain-path settings that label the training scenes (the 1920x1200x33 u16
# workload: limited descriptors, min_variance 2.0, threshold 0.96, step 0.1).
LABEL_CONFIG = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
TRAIN_SEED0 = 0x7A1E0000     # scene seeds TRAIN_SEED0 + i
EVAL_SEED0 = 0xE7A10000      # disjoint from the training seeds


@dataclass
class TrainSettings:
    scenes: int = 2000
    epochs: int = 3
    batch: int = 4
    lr: float = 2e-3
    lr_final: float = 1e-4
    min_valid: int = 8
    seed: int = 0
    seed0: int = TRAIN_SEED0


def scene_labels(spec, seed: int, cfg: Optional[Config] = None, min_valid: int = 8,
                 device="cuda"):
    """(left, right, labels) of one scene: labels from the plain path."""
    cfg = cfg or Config(**LABEL_CONFIG)
    L, R, _ = scene_device(spec, seed, device)
    disp, _ = match_device(L, R, cfg)
    return L, R, label(disp, min_valid)


def train(spec=None, settings: Optional[TrainSettings] = None, opts: Optional[WindowOptions] = None,
          params0: Optional[np.ndarray] = None, device="cuda",
          log: Optional[Callable[[str], None]] = print):
    """Trains the window network on settings.scenes synthetic scenes of `spec`
    (default: 1920x1200x33 u16) labelled by the plain path. Returns
    (params float32 array, history list of (step, mean loss))."""
    torch = _torch()
    spec = spec or _scene.SCENE_1920x1200_U16
    st = settings or TrainSettings()
    opts = opts or WindowOptions()
    t0 = time.time()
    labels = []
    for i in range(st.scenes):
        _, _, lab = scene_labels(spec, st.seed0 + i, min_valid=st.min_valid, device=device)
        labels.append(lab)
    torch.cuda.synchronize()
    if log:
        log(f"labelled {st.scenes} scenes in {time.time() - t0:.1f} s")

    params = params_device(params0 if params0 is not None else init_params(st.seed), device)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    grad = torch.zeros_like(params)
    loss = torch.zeros(2, dtype=torch.float32, device=device)
    rng = np.random.default_rng(st.seed)
    steps_per_epoch = (st.scenes + st.batch - 1) // st.batch
    total = steps_per_epoch * st.epochs
    history = []
    step = 0
    for ep in range(st.epochs):
        order = rng.permutation(st.scenes)
        run_loss, run_cnt = 0.0, 0.0
        for b in range(steps_per_epoch):
            idx = order[b * st.batch:(b + 1) * st.batch]
            grad.zero_()
            loss.zero_()
            for i in idx:
                L, R, _ = scene_device(spec, st.seed0 + int(i), device, frames=CNN_FRAMES)
                loss_grad(L, R, params, labels[i], grad, loss, opts)
            step += 1
            lr = st.lr_final + 0.5 * (st.lr - st.lr_final) * (1 + math.cos(math.pi * (step - 1) / total))
            adam(params, m, v, grad, loss[1:2], lr, step)
            if step % 50 == 0 or step == total:
                l, c = loss.tolist()
                run_loss, run_cnt = l, c
                history.append((step, l / max(c, 1.0)))
                if log:
                    log(f"epoch {ep} step {step}/{total} loss {l / max(c, 1.0):.4f} "
                        f"lr {lr:.2e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return params.cpu().numpy(), history


def window_hits(center, hw, labels):
    """Fraction of labelled cells whose window contains the label (within
    the half-width) and the mean |centre - label| over them."""
    torch = _torch()
    ok = torch.isfinite(labels)
    err = (center - labels).abs()[ok]
    hit = (err <= hw.float()[ok]).float().mean().item() if err.numel() else float("nan")
    return hit, err.mean().item() if err.numel() else float("nan")


def evaluate(params, spec=None, scenes: int = 200, opts: Optional[WindowOptions] = None,
             precision: int = 4, seed0: int = EVAL_SEED0, device="cuda",
             log: Optional[Callable[[str], None]] = print):
    """Learned vs plain path on `scenes` unseen scenes: agreement (share of
    plain-valid pixels that are valid in the learned path and within 0.5 px),
    learned-valid/plain-invalid share, fallback rate and GPU time of both
    paths (events around each call)."""
    torch = _torch()
    from . import Config as _Config
    spec = spec or _scene.SCENE_1920x1200_U16
    opts = opts or WindowOptions()
    params = params_device(params, device)
    cfg_plain = _Config(**LABEL_CONFIG)
    cfg_win = _Config(precision=precision, **LABEL_CONFIG)
    agree, extra, fb, t_plain, t_win, hits = [], [], [], [], [], []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for k in range(scenes):
        L, R, _ = scene_device(spec, seed0 + k, device)
        ev[0].record()
        a, _ = match_device(L, R, cfg_plain)
        ev[1].record()
        b, _, st = match_windowed_device(L, R, cfg_win, params, opts, stats=True)
        ev[2].record()
        torch.cuda.synchronize()
        t_plain.append(ev[0].elapsed_time(ev[1]))
        t_win.append(ev[1].elapsed_time(ev[2]))
        va = torch.isfinite(a)
        vb = torch.isfinite(b)
        ok = vb & ((b - a).abs() <= 0.5)
        agree.append((ok & va).sum().item() / max(va.sum().item(), 1))
        extra.append((vb & ~va).float().mean().item())
        fb.append(st["fallback_pixels"] / st["pixels"])
        if k < 20:
            c, hw, _ = predict(L, R, params, opts)
            hits.append(window_hits(c, hw, label(a))[0])
        if log and (k % 20 == 19 or k == scenes - 1):
            log(f"{k + 1}/{scenes}: agreement min {min(agree):.5f} mean {np.mean(agree):.5f}, "
                f"fallback mean {np.mean(fb):.4f}, plain {np.median(t_plain):.3f} ms, "
                f"learned {np.median(t_win):.3f} ms")
    return {
        "scenes": scenes, "agreement_min": float(min(agree)), "agreement_mean": float(np.mean(agree)),
        "scenes_below_99.5": int(sum(a < 0.995 for a in agree)),
        "learned_valid_plain_invalid": float(np.mean(extra)), "fallback_rate": float(np.mean(fb)),
        "window_hit_rate": float(np.mean(hits)) if hits else float("nan"),
        "plain_ms_median": float(np.median(t_plain)), "learned_ms_median": float(np.median(t_win)),
        "opts": asdict(opts), "precision": precision,
    }


def _main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train / evaluate the learned search window")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=3)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--scenes", type=int, default=200)
    e.add_argument("--precision", type=int, default=4, help="3 MOMENTS, 4 MOMENTS_FAST (default)")
    e.add_argument("--hw-lo", type=int, default=None)
    e.add_argument("--hw-hi", type=int, default=None)
    e.add_argument("--p-conf", type=float, default=None)
    e.add_argument("--fallback", type=int, default=FALLBACK_ALL,
                   help="0 none, 1 all rejected pixels, 2 those in unsure cells")
    e.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        st = TrainSettings(scenes=a.scenes, epochs=a.epochs, batch=a.batch, lr=a.lr)
        t0 = time.time()
        params, hist = train(settings=st, log=lambda m: print(m, flush=True))
        save(a.out, params, {**asdict(st), "seconds": round(time.time() - t0, 1),
                             "final_loss": hist[-1][1] if hist else None})
        print(f"saved {a.out}", flush=True)
    else:
        params, meta = load(a.weights)
        o = WindowOptions()
        if a.hw_lo is not None:
            o.hw_lo = a.hw_lo
        if a.hw_hi is not None:
            o.hw_hi = a.hw_hi
        if a.p_conf is not None:
            o.p_conf = a.p_conf
        o.fallback = a.fallback
        res = evaluate(params, scenes=a.scenes, opts=o, precision=a.precision,
                       log=lambda m: print(m, flush=True))
        print(json.dumps(res), flush=True)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)


if __name__ == "__main__":
    _main()
