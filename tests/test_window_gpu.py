"""Learned search window on the GPU, through the C ABI.

- the device scene generator against the host generator;
- the CNN's loss and gradient (HIP forward/backward kernels) against the
  numpy reference oracle/cnn_ref.py, itself pinned by finite differences;
- Adam against its formula, labels against their definition;
- the windowed matching path: bit-exact to the plain path when the window
  spans the whole row, and the exact fallback invariant (a pixel the
  learned path rejects is rejected by the plain path) for any network;
- the trained network (committed weights): agreement with the plain path on
  unseen 1920x1200x33 scenes and its gain over an untrained network.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BENCH_CFG = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)


def host(t):
    import torch
    torch.cuda.synchronize()
    if t.dtype == torch.uint16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16)
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def W():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from libbicos_amd import window
    return window


# ---------------------------------------------------------------- scenes

@pytest.mark.parametrize("which", ["1920x1200_u16", "640x480_u8"])
def test_scene_device_matches_host(W, which):
    from libbicos_amd import scene
    spec = scene.SCENE_1920x1200_U16 if which == "1920x1200_u16" else scene.SCENE_640x480_U8
    kw = dict(height=40, width=200, frames=5)
    L, R, D = scene.generate(spec, seed=11, threads=4, **kw)
    dL, dR, dD = W.scene_device(spec, 11, frames=kw["frames"], height=kw["height"], width=kw["width"])
    gL, gR, gD = host(dL), host(dR), host(dD)
    # same construction and hashes; device and host libm may round
    # sin/cos/log differently in the last place
    assert np.abs(gD - D).max() <= 1e-3
    for a, b in ((gL, L), (gR, R)):
        diff = np.abs(a.astype(np.int64) - b.astype(np.int64))
        assert (diff == 0).mean() >= 0.999 and diff.max() <= 1


def test_scene_frames_independent_of_count(W):
    from libbicos_amd import scene
    spec = scene.SCENE_1920x1200_U16
    a = W.scene_device(spec, 5, frames=6, height=32, width=96)
    b = W.scene_device(spec, 5, frames=4, height=32, width=96)
    assert np.array_equal(host(a[0])[:4], host(b[0]))
    assert np.array_equal(host(a[1])[:4], host(b[1]))


# ---------------------------------------------------------------- labels

@pytest.mark.parametrize("kind", ["f32", "i16"])
def test_labels(W, kind):
    import torch
    rng = np.random.default_rng(3)
    H, Wd = 37, 50
    d = rng.uniform(0, 100, size=(H, Wd)).astype(np.float32)
    d[rng.random((H, Wd)) < 0.4] = np.nan
    if kind == "i16":
        di = np.where(np.isnan(d), -32768, np.round(d)).astype(np.int16)
        t = torch.from_numpy(di).cuda()
        ref_in = np.where(di == -32768, np.nan, di.astype(np.float32))
    else:
        t = torch.from_numpy(d).cuda()
        ref_in = d
    got = host(W.label(t, min_valid=8))
    Hq, Wq = H // 4, Wd // 4
    blk = ref_in[:Hq * 4, :Wq * 4].reshape(Hq, 4, Wq, 4)
    cnt = np.isfinite(blk).sum(axis=(1, 3))
    s = np.nansum(blk, axis=(1, 3))
    exp = np.where(cnt >= 8, s / np.maximum(cnt, 1), np.nan).astype(np.float32)
    assert np.allclose(got, exp, rtol=1e-6, atol=1e-4, equal_nan=True)


# ---------------------------------------------------------- loss/gradient

def small_scene(W, seed, H=48, Wd=160):
    from libbicos_amd import scene
    spec = scene.SCENE_1920x1200_U16
    L, R, D = W.scene_device(spec, seed, frames=4, height=H, width=Wd)
    Hq, Wq = H // 4, Wd // 4
    gt = host(D)[:Hq * 4, :Wq * 4].reshape(Hq, 4, Wq, 4).mean(axis=(1, 3)).astype(np.float32)
    gt[0, :5] = np.nan           # some unlabelled cells
    return L, R, gt


@pytest.mark.parametrize("bins", [80, 30])
def test_loss_grad_matches_reference(W, bins):
    import torch
    import cnn_ref
    L, R, gt = small_scene(W, 21)
    gt = np.minimum(gt, 4 * (bins - 9))   # keep most labels inside the bins
    p = W.init_params(4)
    p[cnn_ref.b_off(0):cnn_ref.b_off(0) + 32] = np.random.default_rng(1).normal(0, 0.1, 32)
    in_scale = 8 / 65535
    loss_ref, g_ref, _ = cnn_ref.loss_and_grad(host(L), host(R), gt, p.astype(np.float64),
                                               in_scale, bf=True, bins=bins)
    params = W.params_device(p)
    grad = torch.zeros_like(params)
    loss = torch.zeros(2, dtype=torch.float32, device="cuda")
    opts = W.WindowOptions(in_scale=in_scale, bins=bins)
    lab = torch.from_numpy(gt).cuda()
    W.loss_grad(L, R, params, lab, grad, loss, opts)
    l, cnt = host(loss)
    g = host(grad) / max(cnt, 1)
    assert cnt > 0
    assert abs(l / cnt - loss_ref) <= 1e-4 * max(1.0, abs(loss_ref)), (l / cnt, loss_ref)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref)
    assert err.max() <= 2e-3 * scale, (err.max(), scale)
    assert np.linalg.norm(g - g_ref) <= 1e-3 * np.linalg.norm(g_ref)
    # accumulation: a second call adds the same again
    W.loss_grad(L, R, params, lab, grad, loss, opts)
    assert np.allclose(host(grad), 2 * g * cnt, rtol=1e-5, atol=1e-6 * scale * cnt)


def test_adam_matches_formula(W):
    import torch
    rng = np.random.default_rng(0)
    p = rng.normal(size=W.PARAM_COUNT).astype(np.float32)
    m = rng.normal(size=W.PARAM_COUNT).astype(np.float32) * 0.1
    v = rng.random(W.PARAM_COUNT).astype(np.float32) * 0.01
    g = rng.normal(size=W.PARAM_COUNT).astype(np.float32)
    cnt = np.array([4.0], np.float32)
    tp, tm, tv = (torch.from_numpy(a.copy()).cuda() for a in (p, m, v))
    W.adam(tp, tm, tv, torch.from_numpy(g).cuda(), torch.from_numpy(cnt).cuda(), 1e-3, 3)
    b1, b2 = np.float64(np.float32(0.9)), np.float64(np.float32(0.999))
    gs = g.astype(np.float64) / 4.0
    m2 = b1 * m + (1 - b1) * gs
    v2 = b2 * v + (1 - b2) * gs * gs
    p2 = p - 1e-3 * (m2 / (1 - b1 ** 3)) / (np.sqrt(v2 / (1 - b2 ** 3)) + 1e-8)
    assert np.allclose(host(tm), m2, rtol=1e-5, atol=1e-7)
    assert np.allclose(host(tv), v2, rtol=1e-5, atol=1e-9)
    assert np.allclose(host(tp), p2, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(W):
    from libbicos_amd import scene
    spec = scene.SCENE_1920x1200_U16
    import torch
    scenes = []
    for i in range(6):
        L, R, D = W.scene_device(spec, 900 + i, height=96, width=384)
        d, _ = __import__("libbicos_amd").match_device(L, R, __import__("libbicos_amd").Config(**BENCH_CFG))
        scenes.append((L[:4].contiguous(), R[:4].contiguous(), W.label(d)))
    params = W.params_device(W.init_params(0))
    m, v = torch.zeros_like(params), torch.zeros_like(params)
    grad = torch.zeros_like(params)
    loss = torch.zeros(2, dtype=torch.float32, device="cuda")
    opts = W.WindowOptions()

    def eval_loss():
        loss.zero_()
        g = torch.zeros_like(params)
        for L, R, lab in scenes:
            W.loss_grad(L, R, params, lab, g, loss, opts)
        l, c = host(loss)
        return l / c

    first = eval_loss()
    for step in range(1, 41):
        grad.zero_()
        loss.zero_()
        for L, R, lab in scenes[:4]:
            W.loss_grad(L, R, params, lab, grad, loss, opts)
        W.adam(params, m, v, grad, loss[1:2], 3e-3, step)
    last = eval_loss()
    print(f"loss {first:.3f} -> {last:.3f}")
    assert last < 0.8 * first


# ------------------------------------------------------ MFMA inference

@pytest.mark.parametrize("bins", [80, 144])
def test_mfma_network_matches_valu(W, bins):
    """The inference network on the matrix cores computes the training
    kernels' network (BF16 weights and activations, FP32 sums) up to
    summation order: same windows almost everywhere."""
    from libbicos_amd import scene
    L, R, _ = W.scene_device(scene.SCENE_1920x1200_U16, 77, frames=4, height=96, width=640)
    p = W.load()[0] if os.path.exists(W.DEFAULT_WEIGHTS) else W.init_params(5)
    params = W.params_device(p)
    opts = W.WindowOptions(bins=bins)
    try:
        W.set_cnn_impl(W.CNN_VALU)
        cv, hv, pv = (host(t) for t in W.predict(L, R, params, opts))
    finally:
        W.set_cnn_impl(W.CNN_MFMA)
    cm, hm, pm = (host(t) for t in W.predict(L, R, params, opts))
    same = np.abs(cm - cv) <= 0.05
    print(f"bins {bins}: centres equal {same.mean():.4f}, max |dp| {np.abs(pm - pv).max():.4f}")
    assert same.mean() >= 0.98
    assert (hm == hv).mean() >= 0.98
    assert np.median(np.abs(pm - pv)) <= 1e-3


@pytest.mark.parametrize("bins", [80, 144])
def test_fused_network_equals_layers(W, bins):
    """The four layers in one kernel give the per-layer kernels' windows
    bit for bit (same products and rounding; the crop's 25 x 162 cells leave
    partial tiles at the right and bottom edges)."""
    from libbicos_amd import scene
    L, R, _ = W.scene_device(scene.SCENE_1920x1200_U16, 78, frames=4, height=100, width=648)
    p = W.load()[0] if os.path.exists(W.DEFAULT_WEIGHTS) else W.init_params(6)
    params = W.params_device(p)
    opts = W.WindowOptions(bins=bins)
    try:
        W.set_cnn_impl(W.CNN_MFMA_LAYERS)
        ref = [host(t) for t in W.predict(L, R, params, opts)]
    finally:
        W.set_cnn_impl(W.CNN_MFMA)
    got = [host(t) for t in W.predict(L, R, params, opts)]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b, equal_nan=True)


# ------------------------------------------------------ windowed matching

def c1_crop(W, seed, H=64, Wd=480, frames=33):
    from libbicos_amd import scene
    L, R, _ = W.scene_device(scene.SCENE_1920x1200_U16, seed, frames=frames, height=H, width=Wd)
    return L, R


@pytest.mark.parametrize("prec", ["single", "moments", "moments_fast"])
@pytest.mark.parametrize("variant", ["nodup", "cons"])
@pytest.mark.parametrize("subpix", [True, False])
def test_full_width_window_equals_plain(W, variant, subpix, prec):
    """A window spanning the row makes the windowed search the full search:
    the learned path must then reproduce the plain path bit for bit (for
    MOMENTS_FAST without a corrmap: the fused windowed pass, its windows
    longer than the staged columns)."""
    import libbicos_amd as B
    L, R = c1_crop(W, 31)
    kw = dict(BENCH_CFG)
    if not subpix:
        kw.pop("subpixel_step")
    if prec == "moments":
        kw["precision"] = B.Precision.Moments
    if prec == "moments_fast":
        kw["precision"] = B.Precision.MomentsFast
    if variant == "cons":
        kw["variant"] = B.Variant.Consistency(max_lr_diff=1, no_dupes=False)
    cfg = B.Config(**kw)
    params = W.params_device(W.init_params(1))
    opts = W.WindowOptions(hw_lo=4000, hw_hi=4000)
    d_plain, c_plain = B.match_device(L, R, cfg, want_corrmap=True)
    d_win, c_win, st = W.match_windowed_device(L, R, cfg, params, opts, want_corrmap=True, stats=True)
    a, b = host(d_plain), host(d_win)
    assert np.array_equal(a, b, equal_nan=True)
    assert np.array_equal(host(c_plain), host(c_win), equal_nan=True)
    assert st["pixels"] == a.size
    # without a corrmap (the MOMENTS stage proper)
    a = host(B.match_device(L, R, cfg)[0])
    b = host(W.match_windowed_device(L, R, cfg, params, opts)[0])
    assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("impl", ["mfma", "mfma_i8", "valu", "moments", "moments_fast", "moments_fast_valu"])
@pytest.mark.parametrize("variant", ["nodup", "cons"])
@pytest.mark.parametrize("subpix", [True, False])
@pytest.mark.parametrize("hw", [0, 3])
def test_fallback_invariant_untrained(W, variant, subpix, hw, impl):
    """With any network, every pixel the learned path rejects was sent
    through the exact full search and NCC stage: the plain path rejects it
    too. Pixels the learned path accepts either match the plain result or
    pass the NCC threshold at another disparity. Every row-search kernel
    (fp4 and int8 matrix cores, vector ALU) serves the listed pixels."""
    import libbicos_amd as B
    if impl.startswith("moments") and not subpix:
        pytest.skip("the MOMENTS precisions concern the subpixel stage")
    B.set_search_impl({"valu": B.SEARCH_VALU, "mfma_i8": B.SEARCH_MFMA_I8,
                       "moments_fast_valu": B.SEARCH_VALU}.get(impl, B.SEARCH_MFMA))
    try:
        _fallback_invariant(W, B, variant, subpix, hw,
                            {"moments": B.Precision.Moments, "moments_fast": B.Precision.MomentsFast,
                             "moments_fast_valu": B.Precision.MomentsFast}.get(impl))
    finally:
        B.set_search_impl(B.SEARCH_MFMA)


def _fallback_invariant(W, B, variant, subpix, hw, moments):
    L, R = c1_crop(W, 32 + hw)
    kw = dict(BENCH_CFG)
    if not subpix:
        kw.pop("subpixel_step")
    if moments is not None:  # the plain path runs the same precision
        kw["precision"] = moments
    if variant == "cons":
        kw["variant"] = B.Variant.Consistency(max_lr_diff=1, no_dupes=False)
    cfg = B.Config(**kw)
    params = W.params_device(W.init_params(2))
    opts = W.WindowOptions(hw_lo=hw, hw_hi=hw)
    d_plain, _ = B.match_device(L, R, cfg)
    if moments is not None:  # no corrmap: the MOMENTS stages proper
        d_win, c_win, st = W.match_windowed_device(L, R, cfg, params, opts, stats=True)
    else:
        d_win, c_win, st = W.match_windowed_device(L, R, cfg, params, opts, want_corrmap=True, stats=True)
    a, b = host(d_plain), host(d_win)
    inv_a = np.isnan(a) if subpix else a == -32768
    inv_b = np.isnan(b) if subpix else b == -32768
    assert not (inv_b & ~inv_a).any(), "learned path rejected a pixel the plain path accepts"
    if c_win is not None:
        cw = host(c_win)
        assert (cw[~inv_b] >= 0.96 - 1e-6).all()
    assert st["fallback_pixels"] > 0
    print(f"hw {hw}: fallback {st['fallback_pixels'] / a.size:.3f}, "
          f"agree {(a == b)[~inv_a].mean():.4f}")


# ----------------------------------------- windowed path against the oracle

def dev_np(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda().view(torch.uint16)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def oracle_windows(rng, rows, cols, widths=(0, 6, 12)):
    """Cell windows covering the interior, both clipped row ends, empty
    windows (wholly outside the row) and exact half-way centres."""
    Hq, Wq = rows // 4, cols // 4
    center = rng.uniform(-20, 300, size=(Hq, Wq)).astype(np.float32)
    kind = rng.integers(0, 6, size=(Hq, Wq))
    center[kind == 1] = -cols - 30.0
    center[kind == 2] = cols + 30.0
    center[kind == 3] = np.round(center[kind == 3]) + 0.5
    center[kind == 4] = rng.uniform(-15, 15, size=int((kind == 4).sum()))   # clipped at col 0 / W-1
    hw = rng.choice(np.array(widths, np.int16), size=(Hq, Wq)).astype(np.int16)
    return center, hw


@pytest.mark.parametrize("words", [1, 2, 4, 8])
def test_search_windowed_matches_oracle(W, oracle, words):
    """The windowed search kernel against the oracle's window-restricted
    bicos search (bicos.hpp:50-113 restricted): half-widths 0 / 6 / 12,
    windows clipped at both row ends, empty windows, half-way centres, every
    flag set and L/R limit, duplicate-rich descriptors, ragged sizes."""
    rng = np.random.default_rng(500 + words)
    for rows, cols in ((9, 70), (12, 1920), (5, 333)):
        pool = rng.integers(0, 2**32, size=(24, words), dtype=np.uint64).astype(np.uint32)
        d0 = pool[rng.integers(0, 24, size=(rows, cols))]
        d1 = pool[rng.integers(0, 24, size=(rows, cols))]
        center, hw = oracle_windows(rng, rows, cols)
        q0, q1, qc, qh = dev_np(d0), dev_np(d1), dev_np(center), dev_np(hw)
        for flags in (0, 1, 2, 3):
            for lr in (0, 2):
                got = host(W.search_windowed(q0, q1, flags, lr, qc, qh))
                exp = oracle.search_window(d0, d1, flags, lr, center, hw)
                assert np.array_equal(got, exp), (rows, cols, flags, lr)


def _windowed_vs_oracle(W, oracle, L, R, params, opts, cfg_kw, okw, corrmap):
    """bicos_hip_match_windowed against the oracle's windowed match fed with
    the windows the network predicts for the same stacks."""
    import libbicos_amd as B
    c, hw, _ = W.predict(L, R, params, opts)
    center, halfw = host(c), host(hw)
    cfg = B.Config(**cfg_kw)
    d, cm, st = W.match_windowed_device(L, R, cfg, params, opts, want_corrmap=corrmap, stats=True)
    exp, exp_cm, nfb = oracle.match_windowed(host(L), host(R), center, halfw, hw_lo=opts.hw_lo,
                                             fallback=opts.fallback, corrmap=corrmap, **okw)
    got = host(d)
    assert got.dtype == exp.dtype and np.array_equal(got, exp, equal_nan=True)
    if corrmap and exp_cm is not None:
        assert np.array_equal(host(cm), exp_cm, equal_nan=True)
    return got, st, nfb, halfw


@pytest.mark.parametrize("fallback", ["all", "none", "unsure"])
@pytest.mark.parametrize("stage", ["subpixel", "agree", "search_only", "double"])
def test_windowed_match_matches_oracle_c1(W, oracle, fallback, stage):
    """1920-wide c1 crops (33 frames u16, threshold 0.96, min_variance 2,
    step 0.1): the learned path with the committed network's windows (half-
    widths 6 / 12) bit for bit equal to the oracle's windowed search ->
    NCC stage -> fallback full-row search -> NCC stage; the fallback set has
    the oracle's size. Also without subpixel step, without threshold (the
    windowed search alone) and in FP64."""
    import libbicos_amd as B
    params = W.params_device(W.load()[0]) if os.path.exists(W.DEFAULT_WEIGHTS) else W.params_device(W.init_params(9))
    fb = {"all": W.FALLBACK_ALL, "none": W.FALLBACK_NONE, "unsure": W.FALLBACK_UNSURE}[fallback]
    opts = W.WindowOptions(fallback=fb)
    cfg_kw = dict(BENCH_CFG)
    okw = dict(thr=0.96, step=0.1, min_variance=2.0)
    if stage == "agree":
        cfg_kw.pop("subpixel_step"); okw["step"] = None
    elif stage == "search_only":
        cfg_kw = dict(nxcorr_threshold=None); okw = dict(thr=None)
    elif stage == "double":
        cfg_kw["precision"] = B.Precision.Double; okw["double"] = True
    L, R = c1_crop(W, 61, H=48, Wd=1920)
    got, st, nfb, halfw = _windowed_vs_oracle(W, oracle, L, R, params, opts, cfg_kw, okw, corrmap=stage != "search_only")
    assert (halfw == 6).any() and (halfw == 12).any() or not os.path.exists(W.DEFAULT_WEIGHTS)
    if stage != "search_only" and fallback != "none":
        assert st["fallback_pixels"] == nfb > 0
    print(f"{stage} {fallback}: fallback {nfb / got.size:.4f}")


def test_windowed_match_matches_oracle_untrained(W, oracle):
    """An untrained network's windows (centres all over the range, clipped
    and empty windows) with narrow half-widths: the learned path still
    equals the oracle's windowed match bit for bit (fallback ALL: the plain
    output wherever the windowed pass is rejected)."""
    params = W.params_device(W.init_params(3))
    L, R = c1_crop(W, 62, H=32, Wd=640)
    opts = W.WindowOptions(hw_lo=2, hw_hi=5, fallback=W.FALLBACK_ALL)
    _windowed_vs_oracle(W, oracle, L, R, params, opts, dict(BENCH_CFG),
                        dict(thr=0.96, step=0.1, min_variance=2.0), corrmap=True)


def test_windowed_match_matches_oracle_4096_band(W, oracle):
    """The learned form of the 4096x3000x64 stress workload on a 16-row band:
    64 frames (256-bit descriptors), Consistency(max_lr_diff 1), step 0.05,
    threshold 0.96, min_variance 2, K = 144 bins, fallback ALL; bit for bit
    equal to the oracle's windowed match; and every pixel the learned path
    rejects is rejected by the oracle's plain match."""
    import libbicos_amd as B
    from libbicos_amd import scene
    spec = scene.SCENE_4096x3000_U16
    L, R, _ = W.scene_device(spec, 0x600DF00D, height=16)
    params = W.params_device(W.load()[0]) if os.path.exists(W.DEFAULT_WEIGHTS) else W.params_device(W.init_params(9))
    opts = W.WindowOptions(bins=144, fallback=W.FALLBACK_ALL)
    cfg_kw = dict(nxcorr_threshold=0.96, subpixel_step=0.05, min_variance=2.0,
                  variant=B.Variant.Consistency(max_lr_diff=1, no_dupes=False))
    okw = dict(thr=0.96, step=0.05, min_variance=2.0, consistency=(1, False))
    got, st, nfb, _ = _windowed_vs_oracle(W, oracle, L, R, params, opts, cfg_kw, okw, corrmap=False)
    assert got.shape == (16, 4096) and st["fallback_pixels"] == nfb
    plain, _ = oracle.match(host(L), host(R), **okw)
    assert not (np.isnan(got) & np.isfinite(plain)).any()


@pytest.mark.parametrize("variant", ["nodupes", "consistency"])
def test_windowed_match_matches_oracle_u8_three_chunks(W, oracle, variant):
    """12 frames u8 (64-bit descriptors) on 2689-column rows: the fallback
    list search stages 64-bit candidate rows in chunks of 1344 columns, so a
    row is two full chunks and one of a single column. An untrained
    network's narrow windows send most pixels to the fallback; the learned
    path equals the oracle's windowed match bit for bit."""
    import libbicos_amd as B
    from libbicos_amd import scene
    assert B.descriptor_bits(12, 0) == 64
    L, R, _ = W.scene_device(scene.SCENE_640x480_U8, 0xC0FFEE, height=8, width=2689)
    params = W.params_device(W.init_params(5))
    opts = W.WindowOptions(hw_lo=2, hw_hi=5, fallback=W.FALLBACK_ALL)
    cfg_kw = dict(nxcorr_threshold=0.9, subpixel_step=0.25, min_variance=1.0)
    okw = dict(thr=0.9, step=0.25, min_variance=1.0)
    if variant == "consistency":
        cfg_kw["variant"] = B.Variant.Consistency(max_lr_diff=2, no_dupes=True)
        okw["consistency"] = (2, True)
    got, st, nfb, _ = _windowed_vs_oracle(W, oracle, L, R, params, opts, cfg_kw, okw, corrmap=True)
    assert got.shape == (8, 2689) and st["fallback_pixels"] == nfb > got.size // 2


def test_windowed_errors(W):
    import torch
    import libbicos_amd as B
    L, R = c1_crop(W, 40, H=16, Wd=64)
    params = W.params_device(W.init_params(0))
    with pytest.raises(ValueError, match="4 frames"):
        W.match_windowed_device(L[:3], R[:3], B.Config(), params)
    with pytest.raises(B.Exception):
        W.match_windowed_device(L, R, B.Config(), params[:10].contiguous())
    with pytest.raises(ValueError):
        W.match_windowed_device(L, R, B.Config(), params, W.WindowOptions(bins=4))
    with pytest.raises(ValueError):
        W.match_windowed_device(L[:, :3], R[:, :3], B.Config(), params)


# ---------------------------------------------------------- trained weights

def trained(W):
    if not os.path.exists(W.DEFAULT_WEIGHTS):
        pytest.skip("no trained window weights committed")
    p, meta = W.load()
    return W.params_device(p), meta


def test_trained_windows_contain_labels(W):
    """The gain of the trained network over the untrained one: the share of
    labelled cells whose window contains the plain-path disparity."""
    from libbicos_amd import scene
    params, _ = trained(W)
    L, R, lab = W.scene_labels(scene.SCENE_1920x1200_U16, W.EVAL_SEED0 + 7)
    c, hw, _ = W.predict(L, R, params)
    hit, err = W.window_hits(c, hw, lab)
    c0, hw0, _ = W.predict(L, R, W.params_device(W.init_params(0)))
    hit0, err0 = W.window_hits(c0, hw0, lab)
    print(f"trained: hit {hit:.4f} mean err {err:.2f} px; untrained: hit {hit0:.4f} err {err0:.2f} px")
    assert hit >= 0.97 and hit0 <= 0.5


@pytest.mark.parametrize("prec", ["moments_fast", "moments", "half_mfma"])
def test_trained_agreement_full_size_1920x1200x33(W, prec):
    """1920x1200x33 u16, unseen scenes: >= 99.5 % of the plain-valid pixels
    are valid in the learned path and within 0.5 px. MOMENTS_FAST with
    fallback ALL is the configuration the benchmark reports."""
    import libbicos_amd as B
    from libbicos_amd import scene
    params, _ = trained(W)
    cfg = B.Config(precision={"moments_fast": B.Precision.MomentsFast, "moments": B.Precision.Moments,
                              "half_mfma": B.Precision.HalfMFMA}[prec], **BENCH_CFG)
    for k in range(2):
        L, R, _ = W.scene_device(scene.SCENE_1920x1200_U16, W.EVAL_SEED0 + 100 + k)
        a = host(B.match_device(L, R, B.Config(**BENCH_CFG))[0])
        d, _, st = W.match_windowed_device(L, R, cfg, params, W.WindowOptions(fallback=W.FALLBACK_ALL),
                                           stats=True)
        b = host(d)
        va = np.isfinite(a)
        ok = np.isfinite(b) & (np.abs(b - a) <= 0.5)
        agree = ok[va].mean()
        extra = (np.isfinite(b) & ~va).mean()
        fb = st["fallback_pixels"] / st["pixels"]
        print(f"scene {k}: agreement {agree:.5f}, learned-valid/plain-invalid {extra:.5f}, "
              f"fallback {fb:.4f}")
        assert agree >= 0.995
