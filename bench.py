"""Benchmark of BICOS::match on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one BICOS::match call over one synthetic 1920x1200, 33-frame uint16
stereo pair (LIMITED descriptors, min_variance 2.0, nxcorr_threshold 0.96,
subpixel_step 0.1, NoDuplicates) whose stacks are already resident in HBM.
Prints one JSON line: throughput in matched Mpixel/s (`value`: the plain,
bit-exact path), the roofline position of its dominant kernel (per-kernel times
from HIP events recorded by the library on the launch stream during the timed
steps), the learned-window path on the same pair (throughput, ratio to the
plain path, agreement with it, fallback rate; for each full-row fallback
mode), and the CPU oracle timed on the host on the same workload. Steps are
timed without per-kernel events; per-kernel times come from the same number
of further steps with HIP events around every kernel.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "matched Mpixel/s @1920x1200x33 u16; learned vs plain at >=99.5% px agree"

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
INT8_DENSE_TOPS = 5000.0
FP4_DENSE_TFLOPS = 10100.0
FP32_TFLOPS = 157.3
SUBPIXEL_OPS_PER_SAMPLE = 13   # interpolate, round, cast, centre, two fused steps, ... per grid point and frame
VALU_SIMDS = 1024        # 256 CUs x 4 SIMDs; a wave64 VALU instruction takes 4 cycles
CLOCK_HZ = 2.4e9


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    # untimed steps; the GPU's clocks take a few dozen calls to settle
    p.add_argument("--warmup", type=int, default=30)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--seed", type=int, default=0x600DF00D)
    p.add_argument("--search", choices=("mfma", "mfma_i8", "valu"), default="mfma",
                   help="row-search kernel: int8 matrix cores (default) or vector ALU")
    p.add_argument("--workload", choices=("1920x1200x33", "4096x3000x64", "kernels"),
                   default="1920x1200x33",
                   help="1920x1200x33: the benchmark line (default); 4096x3000x64: the "
                        "Consistency stress workload, plain and learned, with peak device "
                        "memory; kernels: the stages at the reference's kernel benchmark "
                        "shapes (3300x2200 random inputs) beside its RTX 4090 times")
    return p.parse_args()


def kernel_roofline(name: str, avg_ms: float, W: int, H: int, n: int, words: int, fp4: bool = False,
                    npts: int = 20):
    """Algorithmic work of one launch of `name` against the MI355X roof.

    search: W*W*H candidates x (descriptor bits) multiply-accumulates (2 ops
    each) against the dense matrix peak of the operand type the kernel used
    (fp4: 10.1 PFLOP/s; int8: 5.0 POP/s); subpixel stages: W*H pixels x
    npts grid points x n frames x 13 operations (the reference's scan: the
    parabola, rint, the cast, centring and the two fused steps of the NCC
    chains, agree.hpp:143-171) against the FP32 peak (157.3 TFLOP/s);
    transform / agree: bytes the algorithm must move (samples read, outputs
    written) against HBM bandwidth. Where rocprofv3 SQ counters for the
    kernel are committed (profiles/pmc_traffic.txt), `valu_issue` adds the
    vector-ALU issue utilisation: instructions per wave x waves x 4 cycles
    over 1024 SIMDs at 2.4 GHz, over the launch time."""
    s = avg_ms * 1e-3
    extra = {}
    valu = committed_valu(name)
    if valu is not None:
        per_wave, waves = valu
        extra["valu_issue"] = round(per_wave * waves * 4 / (VALU_SIMDS * CLOCK_HZ) / s, 4)
    if name.startswith("search"):
        ops = 2.0 * W * W * H * 32 * words
        peak = FP4_DENSE_TFLOPS if fp4 else INT8_DENSE_TOPS
        return {"bound": "mfma", "achieved": ops / s / 1e12, "peak": peak,
                "unit": "TFLOP/s", "frac": ops / s / 1e12 / peak,
                "work_per_launch": ops, "work_unit": "fp4 MACs x 2" if fp4 else "int8 MACs x 2", **extra}
    if name.startswith("agree_subpixel"):
        ops = float(W) * H * npts * n * SUBPIXEL_OPS_PER_SAMPLE
        return {"bound": "valu", "achieved": ops / s / 1e12, "peak": FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": ops / s / 1e12 / FP32_TFLOPS, "work_per_launch": ops,
                "work_unit": f"pixels x {npts} grid points x {n} frames x {SUBPIXEL_OPS_PER_SAMPLE} FP32 ops",
                **extra}
    if name.startswith("transform"):
        b = 2 * W * H * (n * 2 + 4 * words)  # one launch transforms both views
    else:
        b = W * H * (2 + 2 * n * 2 + 2)
    return {"bound": "hbm", "achieved": b / s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": b / s / 1e9 / HBM_PEAK_GBS, "work_per_launch": b, "work_unit": "bytes", **extra}


def committed_traffic(kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy-kernel calibration), recorded in profiles/."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\s+(\S+)\s+(\d+(?:\.\d+)?)", line)
        if m and kernel.startswith(m.group(1)):
            return float(m.group(2))
    return None


def committed_valu(kernel: str):
    """(VALU instructions per wave, waves per launch) from rocprofv3 SQ
    counters, recorded in profiles/pmc_traffic.txt."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"#\s*(\S+)\s+fetch.*_VALU=(\d+).*waves\s+(\d+)", line)
        if m and kernel == m.group(1):
            return float(m.group(2)), float(m.group(3))
    return None


def cpu_baseline(L, R, cfg_kw):
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU oracle, timed as the baseline
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    t0 = time.perf_counter()
    O.match(L, R, **cfg_kw)
    dt = time.perf_counter() - t0
    n, H, W = L.shape
    # one thread on a band of 100 rows (every stage is row-local)
    O.set_threads(1)
    band = 100
    t0 = time.perf_counter()
    O.match(np.ascontiguousarray(L[:, :band]), np.ascontiguousarray(R[:, :band]), **cfg_kw)
    dt1 = time.perf_counter() - t0
    O.set_threads(threads)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    # the reference's CPU configuration: 640x480 u8, 12 frames, FULL
    # descriptors, threshold 0.9, no subpixel, no min_variance, NoDuplicates
    from libbicos_amd import scene
    L0, R0, _ = scene.generate(scene.SCENE_640x480_U8, seed=0x600DF00D)
    c0_kw = dict(thr=0.9, step=None, min_variance=None, full=True)
    O.match(L0, R0, **c0_kw)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        O.match(L0, R0, **c0_kw)
    dc0 = (time.perf_counter() - t0) / reps
    O.set_threads(1)
    t0 = time.perf_counter()
    O.match(L0, R0, **c0_kw)
    dc01 = time.perf_counter() - t0
    O.set_threads(threads)
    return {"value": W * H / dt / 1e6, "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"the full {W}x{H}x{n} workload once ({dt:.2f} s wall, "
                      f"{threads} threads, oracle/bicos_oracle.cpp)",
            "single_thread": {"value": W * band / dt1 / 1e6, "unit": "Mpixel/s",
                              "sample": f"rows 0..{band - 1} of the same pair ({dt1:.2f} s wall)"},
            "c0_640x480x12_u8_full": {
                "value": 640 * 480 / dc0 / 1e6, "unit": "Mpixel/s", "cores": threads,
                "ms": round(dc0 * 1e3, 2), "single_thread_ms": round(dc01 * 1e3, 2),
                "sample": f"BICOS::match on the 640x480 u8 12-frame pair (FULL, threshold 0.9, "
                          f"no subpixel), mean of {reps} runs on {threads} threads; one run on 1 thread"},
            "cpu_model": model}


def pointcloud_timing(B, disp, args):
    """Point cloud of the plain path's disparity (the reference CLI's --q-store
    path): reprojection with the Q of a rectified rig, then the selection
    save_pointcloud applies; the selection returns its counts, so each call
    ends with a stream synchronisation."""
    import torch
    H, W = disp.shape
    f, B_m = 1400.0, 0.1
    Q = [[1, 0, 0, -(W - 1) / 2], [0, 1, 0, -(H - 1) / 2], [0, 0, 0, f], [0, 0, 1 / B_m, 0]]
    B.reproject(disp, Q)
    B.pointcloud(disp, Q)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        B.reproject(disp, Q)
    e1.record()
    torch.cuda.synchronize()
    rep_ms = e0.elapsed_time(e1) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, counts = B.pointcloud(disp, Q)
    pc_ms = (time.perf_counter() - t0) / args.steps * 1e3
    return {"reproject_ms": round(rep_ms, 4), "reproject_plus_select_ms": round(pc_ms, 4),
            "points": counts["written"],
            "reproject_GBps": round(W * H * (4 + 12) / (rep_ms * 1e-3) / 1e9, 1)}


def learned_path(B, dL, dR, plain_disp, args, W, H, plain_ms):
    """The learned-window path (MOMENTS_FAST subpixel stage, committed
    weights) on the same resident pair: timed like the plain path, then
    compared with the plain disparity (share of plain-valid pixels valid and
    within 0.5 px); fallback ALL is the mode the target is stated for. ALL
    with the MOMENTS stage (reference accept/reject decisions) is timed
    beside it."""
    import torch
    from libbicos_amd import window as Wn
    if not os.path.exists(Wn.DEFAULT_WEIGHTS):
        return None
    params = Wn.params_device(Wn.load()[0])
    base = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    fast = B.Config(precision=B.Precision.MomentsFast, **base)
    mom = B.Config(precision=B.Precision.Moments, **base)
    stream = torch.cuda.current_stream()
    a = plain_disp
    va = torch.isfinite(a)
    res = {"weights": os.path.relpath(Wn.DEFAULT_WEIGHTS, ROOT), "precision": "MOMENTS_FAST",
           "agreement_rule": "plain-valid pixels valid and within 0.5 px of plain"}
    for name, fb, cfg in (("fallback_all", Wn.FALLBACK_ALL, fast), ("fallback_unsure", Wn.FALLBACK_UNSURE, fast),
                          ("fallback_none", Wn.FALLBACK_NONE, fast),
                          ("fallback_all_moments", Wn.FALLBACK_ALL, mom)):
        opts = Wn.WindowOptions(fallback=fb)
        out, _, st = Wn.match_windowed_device(dL, dR, cfg, params, opts, stats=True, stream=stream)
        for _ in range(args.warmup):
            out, _, _ = Wn.match_windowed_device(dL, dR, cfg, params, opts, disparity=out, stream=stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            out, _, _ = Wn.match_windowed_device(dL, dR, cfg, params, opts, disparity=out, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        B.timing_enable(True)
        for _ in range(args.steps):
            out, _, _ = Wn.match_windowed_device(dL, dR, cfg, params, opts, disparity=out, stream=stream)
        torch.cuda.synchronize()
        kst = B.timing_read()
        B.timing_enable(False)
        b = out
        ok = torch.isfinite(b) & ((b - a).abs() <= 0.5)
        res[name] = {
            "value": W * H / (ms * 1e-3) / 1e6, "unit": "Mpixel/s", "ms_per_step": ms,
            "vs_plain": plain_ms / ms,
            "agreement": (ok & va).sum().item() / max(1, va.sum().item()),
            "learned_valid_plain_invalid": (torch.isfinite(b) & ~va).float().mean().item(),
            "fallback_rate": st["fallback_pixels"] / st["pixels"],
            "kernels_ms": {k: round(t / args.steps, 4) for k, (c, t) in sorted(kst.items(), key=lambda kv: -kv[1][1])},
        }
    return res


def precision_modes(B, L, R, dL, dR, args, W, H):
    """c2: the subpixel precision modes on the benchmark pair against the FP64
    oracle (Precision::DOUBLE semantics, agree.cuh:35-65; the oracle on the
    host's cores): max / mean |dNCC| (correlation map; HALF_MFMA rescored
    winners, the exact stage's for the others), max / mean |d disparity|
    over pixels valid in both, the share of pixels whose accept/reject
    decision differs, and ms per match (inputs resident, no corrmap)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    base = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    exp, exp_cm = O.match(L, R, thr=0.96, step=0.1, min_variance=2.0, double=True, corrmap=True)
    fe = np.isfinite(exp)
    out = {"reference": "FP64 oracle on the same pair"}
    for name, prec in (("SINGLE", B.Precision.Single), ("HALF_MFMA", B.Precision.HalfMFMA),
                       ("MOMENTS", B.Precision.Moments), ("MOMENTS_FAST", B.Precision.MomentsFast)):
        cfg = B.Config(precision=prec, **base)
        d, _ = B.match_device(dL, dR, cfg)
        for _ in range(5):
            d, _ = B.match_device(dL, dR, cfg, disparity=d)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            d, _ = B.match_device(dL, dR, cfg, disparity=d)
        e1.record()
        torch.cuda.synchronize()
        got = d.cpu().numpy()
        fg = np.isfinite(got)
        both = fg & fe
        dd = np.abs(got[both].astype(np.float64) - exp[both])
        rec = {"ms": round(e0.elapsed_time(e1) / args.steps, 4),
               "d_max": float(dd.max()), "d_mean": float(dd.mean()),
               "flip_frac": float((fg != fe).mean())}
        if name in ("SINGLE", "HALF_MFMA"):
            # the correlation map of this precision (HALF_MFMA: its rescored winners)
            _, cm = B.match_device(dL, dR, cfg, want_corrmap=True)
            cm = cm.cpu().numpy()
            ok = np.isfinite(cm) & np.isfinite(exp_cm)
            dn = np.abs(cm[ok].astype(np.float64) - exp_cm[ok])
            rec["ncc_max"] = float(dn.max())
            rec["ncc_mean"] = float(dn.mean())
        out[name] = rec
    return out


# The reference's published kernel times on an RTX 4090 at 3300x2200 random
# inputs (bench/baselines/cuda-rtx4090.txt; bench/cuda.cu: one view per
# transform launch at n = the width's largest stack, max_lr_diff 3, agree at
# n = 10 in FP64 with threshold 0.9 and min_variance*n = 10), in ms; the
# search rows are the faster of its two kernels.
RTX4090_MS = {
    "transform u16 limited u32 (n 9)": 0.2016, "transform u16 limited u64 (n 17)": 0.3786,
    "transform u16 limited u128 (n 33)": 1.0756, "transform u8 full u32 (n 6)": 0.2028,
    "transform u8 full u64 (n 8)": 0.3695, "transform u8 full u128 (n 12)": 0.9292,
    "search u32 nodup": 9.4276, "search u64 nodup": 11.3053, "search u128 nodup": 18.8214,
    "search u32 consistency": 11.1487, "search u64 consistency": 18.3765,
    "search u128 consistency": 36.7618, "search u128 nodup+consistency": 18.9464,
    "agree u16 n10 fp64": 1.9494, "agree u8 n10 fp64": 1.9478,
}


def kernel_benchmarks(args):
    """Each stage through its C-ABI entry point at the reference's kernel
    benchmark shapes: kernel time per call over 20 calls (the library's HIP
    events around its launches), beside the reference's RTX 4090 time. The subpixel stage
    has no comparable reference number (its bench passes min_variance and
    step swapped, bench/cuda.cu:175); it is timed at n = 10, step 0.25."""
    import torch
    sys.path.insert(0, ROOT)
    import libbicos_amd as B

    H, W = 2200, 3300
    g = torch.Generator(device="cuda").manual_seed(args.seed)

    def rand_stack(n, dtype):
        hi = 256 if dtype == "u8" else 65536
        t = torch.randint(0, hi, (n, H, W), dtype=torch.int32, device="cuda", generator=g)
        return t.to(torch.uint8) if dtype == "u8" else t.to(torch.int16).view(torch.uint16)

    def timed(fn, calls=20):
        # kernel time per call: the library's HIP events around each launch
        # (all kernels a call launches), after 5 untimed calls
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        B.timing_enable(True)
        for _ in range(calls):
            fn()
        torch.cuda.synchronize()
        st = B.timing_read()
        B.timing_enable(False)
        return sum(total for _, total in st.values()) / calls

    rows = {}
    for dtype, full, n, words in (("u16", False, 9, 1), ("u16", False, 17, 2), ("u16", False, 33, 4),
                                  ("u8", True, 6, 1), ("u8", True, 8, 2), ("u8", True, 12, 4)):
        st = rand_stack(n, dtype)
        out = torch.empty((H, W, words), dtype=torch.int32, device="cuda")
        name = f"transform {dtype} {'full' if full else 'limited'} u{32 * words} (n {n})"
        rows[name] = timed(lambda: B.transform(st, int(full), out=out))
        del st
    for words in (1, 2, 4):
        d0 = torch.randint(-2**31, 2**31 - 1, (H, W, words), dtype=torch.int32, device="cuda", generator=g)
        d1 = torch.randint(-2**31, 2**31 - 1, (H, W, words), dtype=torch.int32, device="cuda", generator=g)
        out = torch.empty((H, W), dtype=torch.int16, device="cuda")
        for label, flags in (("nodup", 1), ("consistency", 2), ("nodup+consistency", 3)):
            if label == "nodup+consistency" and words != 4:
                continue
            rows[f"search u{32 * words} {label}"] = timed(lambda: B.search(d0, d1, flags, 3, out=out))
        del d0, d1
    for dtype in ("u16", "u8"):
        s0, s1 = rand_stack(10, dtype), rand_stack(10, dtype)
        disp = torch.randint(-1, W, (H, W), dtype=torch.int32, device="cuda", generator=g).to(torch.int16)
        work = torch.empty_like(disp)

        def agree():
            work.copy_(disp)  # agree is in place on the disparity
            B.agree(work, s0, s1, 0.9, 1.0, B.Precision.Double)
        rows[f"agree {dtype} n10 fp64"] = timed(agree)
        sub = torch.empty((H, W), dtype=torch.float32, device="cuda")
        rows[f"agree_subpixel {dtype} n10 step0.25 fp32"] = timed(
            lambda: B.agree_subpixel(disp, s0, s1, 0.9, 0.25, 1.0, B.Precision.Single, out=sub))
        rows[f"agree_subpixel {dtype} n10 step0.25 fp64"] = timed(
            lambda: B.agree_subpixel(disp, s0, s1, 0.9, 0.25, 1.0, B.Precision.Double, out=sub))
        del s0, s1
    table = {k: {"ms": round(v, 4), "rtx4090_ms": RTX4090_MS.get(k),
                 "speedup": round(RTX4090_MS[k] / v, 2) if k in RTX4090_MS else None}
             for k, v in rows.items()}

    # Whole match at the size of the reference README's dataset run (33+33
    # frames, LIMITED, threshold 0.96, min_variance 2.0, step 0.1, NoDuplicates:
    # about 44 ms on an RTX 4090, README.md:71,81; 3208x2200 is the size of its
    # example disparity), on a synthetic scene of the 1920x1200x33 generator
    # scaled to that size; inputs resident on the device.
    from libbicos_amd import scene, window as Wn
    spec = scene.SCENE_1920x1200_U16
    Lh, Rh, _ = Wn.scene_device(spec, args.seed, height=2200, width=3208)
    cfg = B.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    d, _ = B.match_device(Lh, Rh, cfg)
    for _ in range(10):
        d, _ = B.match_device(Lh, Rh, cfg, disparity=d)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        d, _ = B.match_device(Lh, Rh, cfg, disparity=d)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    whole = {"shape": "3208x2200x33 u16 (synthetic)", "ms": round(ms, 3),
             "Mpixel/s": round(3208 * 2200 / ms / 1e3, 1), "rtx4090_ms_readme": 44.0,
             "speedup": round(44.0 / ms, 2)}
    print(json.dumps({"workload": "stages at 3300x2200 random inputs (reference kernel benchmarks)",
                      "kernels": table, "match_readme_shape": whole}))


def stress(args):
    """4096x3000, 64 frames/side u16 (256-bit descriptors), Consistency
    (max_lr_diff 1), subpixel step 0.05, threshold 0.96, min_variance 2.0:
    plain and learned paths, device memory in use after each (the library
    keeps its workspace), per-kernel times and algorithmic GB/s."""
    import torch
    sys.path.insert(0, ROOT)
    import libbicos_amd as B
    from libbicos_amd import scene, window as Wn
    spec = scene.SCENE_4096x3000_U16
    W, H, n = spec.width, spec.height, spec.frames
    free0, total = torch.cuda.mem_get_info()
    L, R, _ = Wn.scene_device(spec, args.seed)
    torch.cuda.synchronize()
    B.release_workspace()
    free1, _ = torch.cuda.mem_get_info()
    kw = dict(nxcorr_threshold=0.96, subpixel_step=0.05, min_variance=2.0,
              variant=B.Variant.Consistency(max_lr_diff=1, no_dupes=False))
    out = {"workload": f"BICOS::match {W}x{H}x{n} u16, LIMITED (256-bit), Consistency "
                       "max_lr_diff 1, subpixel_step 0.05, nxcorr_threshold 0.96, min_variance 2.0",
           "stacks_bytes": free0 - free1}
    runs = [("plain", None, B.Config(**kw)),
            ("learned_fallback_all", True, B.Config(precision=B.Precision.MomentsFast, **kw)),
            ("learned_fallback_none", False, B.Config(precision=B.Precision.MomentsFast, **kw))]
    params = Wn.params_device(Wn.load()[0]) if os.path.exists(Wn.DEFAULT_WEIGHTS) else None
    plain = None
    for name, fb, cfg in runs:
        if fb is not None and params is None:
            continue
        opts = Wn.WindowOptions(bins=144, fallback=fb) if fb is not None else None
        call = (lambda: B.match_device(L, R, cfg)[0]) if fb is None else \
            (lambda: Wn.match_windowed_device(L, R, cfg, params, opts)[0])
        d = call()
        torch.cuda.synchronize()
        used = total - torch.cuda.mem_get_info()[0]
        B.timing_enable(True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        steps = max(1, min(args.steps, 5))
        e0.record()
        for _ in range(steps):
            d = call()
        e1.record()
        torch.cuda.synchronize()
        kst = B.timing_read()
        B.timing_enable(False)
        ms = e0.elapsed_time(e1) / steps
        rec = {"ms": ms, "Mpixel/s": W * H / ms / 1e3, "device_memory_in_use_bytes": used,
               "valid_fraction": torch.isfinite(d).float().mean().item(),
               "kernels_ms": {k: round(t / steps, 4) for k, (c, t) in sorted(kst.items(), key=lambda kv: -kv[1][1])}}
        for k, (c, t) in kst.items():
            if k == "transform":
                rec["transform_GBps"] = W * H * 2 * (2 * n + 32) / (t / steps * 1e-3) / 1e9
        if fb is None:
            plain = d
        else:
            va = torch.isfinite(plain)
            ok = torch.isfinite(d) & ((d - plain).abs() <= 0.5)
            rec["agreement"] = (ok & va).sum().item() / max(1, va.sum().item())
            rec["vs_plain"] = out["plain"]["ms"] / ms
        out[name] = rec
        B.release_workspace()
    print(json.dumps(out), flush=True)


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the stages are "
              "row-local, so N GPUs would each take a band of rows)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "4096x3000x64":
        stress(args)
        return
    if args.workload == "kernels":
        kernel_benchmarks(args)
        return

    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    import libbicos_amd as B
    from libbicos_amd import scene

    spec = scene.SCENE_1920x1200_U16
    W, H, n = spec.width, spec.height, spec.frames
    L, R, _ = scene.generate(spec, seed=args.seed)
    cfg = B.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0,
                   mode=B.TransformMode.Limited, variant=B.Variant.NoDuplicates())
    words = B.descriptor_bits(n, cfg.mode) // 32
    B.set_search_impl({"mfma": B.SEARCH_MFMA, "mfma_i8": B.SEARCH_MFMA_I8, "valu": B.SEARCH_VALU}[args.search])
    # the default matrix-core search takes fp4 operands for 64- to 256-bit
    # descriptors on rows up to 4096 columns
    fp4 = args.search == "mfma" and words >= 2 and W <= 4096

    dL = torch.from_numpy(L.view(np.int16)).cuda().view(torch.uint16)
    dR = torch.from_numpy(R.view(np.int16)).cuda().view(torch.uint16)
    stream = torch.cuda.current_stream()
    disp, corr = B.match_device(dL, dR, cfg, stream=stream)
    for _ in range(args.warmup):
        disp, corr = B.match_device(dL, dR, cfg, disparity=disp, stream=stream)
    torch.cuda.synchronize()

    # the timed steps: no per-kernel events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        disp, corr = B.match_device(dL, dR, cfg, disparity=disp, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    # the same steps again with HIP events around every kernel (per-kernel times)
    B.timing_enable(True)
    for _ in range(args.steps):
        disp, corr = B.match_device(dL, dR, cfg, disparity=disp, stream=stream)
    torch.cuda.synchronize()
    kstats = B.timing_read()
    B.timing_enable(False)

    kernels = {k: {"launches": int(c), "avg_ms": t / c} for k, (c, t) in kstats.items()}
    per_step = {k: v["avg_ms"] * v["launches"] / args.steps for k, v in kernels.items()}
    dom = max(per_step, key=per_step.get)
    roof = kernel_roofline(dom, kernels[dom]["avg_ms"], W, H, n, words, fp4)
    roof["kernel"] = dom
    roof["traffic"] = committed_traffic(dom)
    for k, v in kernels.items():
        r = kernel_roofline(k, v["avg_ms"], W, H, n, words, fp4)
        v["bound"] = r["bound"]
        v["roofline_frac"] = round(r["frac"], 4)
        if "valu_issue" in r:
            v["valu_issue"] = r["valu_issue"]

    d = disp.cpu().numpy()
    valid = float(np.isfinite(d).mean())
    learned = learned_path(B, dL, dR, disp, args, W, H, ms_per_step)
    points = pointcloud_timing(B, disp, args)
    precision = None if args.no_cpu_baseline else precision_modes(B, L, R, dL, dR, args, W, H)

    out = {
        "metric": METRIC,
        "value": W * H / (ms_per_step * 1e-3) / 1e6,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("u16 samples, fp4 Hamming products (exact f32 sums), f32 NCC" if fp4 else
                  "u16 samples, u32 Hamming, f32 NCC" if args.search == "valu" else
                  "u16 samples, int8 Hamming products (i32 sums), f32 NCC"),
        "data": "synthetic",
        "config": {"workload": f"plain BICOS::match {W}x{H}x{n} u16, LIMITED, min_variance 2.0, "
                               "nxcorr_threshold 0.96, subpixel_step 0.1, NoDuplicates",
                   "width": W, "height": H, "frames": n, "descriptor_bits": 32 * words,
                   "search_kernel": args.search,
                   "valid_fraction": round(valid, 4), "wall_s_timed": round(wall, 4)},
        "roofline": roof,
        "kernels": kernels,
        "learned": learned,
        "precision_c2": precision,
        "pointcloud": points,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(L, R, dict(thr=0.96, step=0.1, min_variance=2.0,
                                                      full=False))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
