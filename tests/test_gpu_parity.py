"""HIP path vs the CPU oracle, through the C ABI, on the same seeded inputs.

Mirrors the reference's differential tests (test/descriptor_transform.cu,
test/bicos.cu, test/agree.cu, test/integration.cu): random inputs of random
ragged sizes, every descriptor width, input depth, search variant and
precision; plus whole-path runs on synthetic scenes of the benchmark
workloads. The bar is bit-exact equality for descriptors, integer and float
disparities and the correlation maps (NaN-aware); the FP32 NCC/subpixel
tolerance of 1e-5 / 1e-3 px is asserted on top where it applies.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda().view(torch.uint16)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def host(t, dtype=None):
    import torch
    torch.cuda.synchronize()
    if t.dtype == torch.uint16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16)
    a = t.cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def same(a, b):
    """bitwise equality, NaN == NaN"""
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)
    return np.array_equal(a, b)


def rand_stack(rng, n, rows, cols, dtype, levels=None):
    hi = levels if levels else np.iinfo(dtype).max + 1
    return rng.integers(0, hi, size=(n, rows, cols)).astype(dtype)


# ------------------------------------------------------------ transform

LIMITED_N = [2, 3, 4, 5, 8, 9, 10, 16, 17, 18, 25, 32, 33, 34, 48, 64, 65]
FULL_N = list(range(2, 17))


def transform_kernels(gpu, fn):
    """Runs fn with the library's per-kernel events on; the kernel names run."""
    gpu.timing_enable(True)
    try:
        out = fn()
        st = gpu.timing_read()
    finally:
        gpu.timing_enable(False)
    return out, set(st)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("aligned", [True, False])
def test_transform_bit_exact(gpu, oracle, dtype, full, aligned):
    """Every n of either mode, at widths that are a multiple of 4 samples
    (the 4-pixel-per-lane kernels, the production path; 2 pixels per lane for
    u16 stacks deeper than 33 frames) and at odd widths (one pixel per lane);
    the kernel that ran is checked by name."""
    rng = np.random.default_rng(100 + full + 2 * (dtype == np.uint16) + 4 * aligned)
    for n in (FULL_N if full else LIMITED_N):
        rows = int(rng.integers(8, 40))
        cols = 4 * int(rng.integers(64, 257)) if aligned else 2 * int(rng.integers(128, 514)) + 1
        for levels in (None, 3):          # 3 levels: equal samples and equal sums are common
            s = rand_stack(rng, n, rows, cols, dtype, levels)
            exp = oracle.transform(s, full=full)
            q = dev(s)
            got, ran = transform_kernels(gpu, lambda: host(gpu.transform(q, int(full)), np.uint32))
            assert ran == ({"transform"} if aligned else {"transform_scalar"}), (n, cols, ran)
            assert np.array_equal(got, exp), (n, levels, cols)


def test_transform_deep_u16_aligned(gpu, oracle):
    """16-bit stacks deeper than 33 frames at row-aligned widths: the
    2-pixel-per-lane limited kernel (256-bit descriptors)."""
    rng = np.random.default_rng(7)
    for n, cols in ((34, 1028), (48, 772), (64, 4096), (65, 516)):
        for levels in (None, 3):
            s = rand_stack(rng, n, 12, cols, np.uint16, levels)
            exp = oracle.transform(s, full=False)
            got = host(gpu.transform(dev(s), 0), np.uint32)
            assert np.array_equal(got, exp), (n, cols, levels)


def test_transform_wider_than_needed(gpu, oracle):
    rng = np.random.default_rng(3)
    s = rand_stack(rng, 9, 16, 300, np.uint16)
    for words in (1, 2, 4, 8):
        got = host(gpu.transform(dev(s), 0, words=words), np.uint32)
        assert np.array_equal(got, oracle.transform(s, full=False, words=words))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_transform_pitched_and_unaligned_frames(gpu, oracle, dtype):
    """frames with a row pitch larger than the width (GpuMat-style padding),
    and frames starting one sample off alignment (the one-pixel-per-lane
    kernels), in both modes"""
    rng = np.random.default_rng(4)
    big = rand_stack(rng, 12, 40, 512, dtype)
    for off in (0, 1):
        s = np.ascontiguousarray(big[:, :, off:off + 333])
        d = dev(big)[:, :, off:off + 333]
        for full in (False, True):
            got = host(gpu.transform(d, int(full)), np.uint32)
            assert np.array_equal(got, oracle.transform(s, full=full)), (off, full)


# --------------------------------------------------------------- search

@pytest.fixture(params=["mfma", "mfma_i8", "valu"])
def search_impl(request, gpu):
    """every row-search kernel: matrix cores (default: fp4 products for
    descriptors of 64 bits and more, int8 for 32-bit ones), int8 matrix
    cores throughout, and vector ALU"""
    gpu.set_search_impl({"mfma": gpu.SEARCH_MFMA, "mfma_i8": gpu.SEARCH_MFMA_I8,
                         "valu": gpu.SEARCH_VALU}[request.param])
    yield request.param
    gpu.set_search_impl(gpu.SEARCH_MFMA)


@pytest.mark.parametrize("words", [1, 2, 4, 8])
@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_search_exact(gpu, oracle, search_impl, words, flags):
    """test/bicos.cu: random descriptors, max_lr_diff 3, exact equality."""
    rng = np.random.default_rng(1000 + 10 * words + flags)
    rows, cols = int(rng.integers(16, 48)), int(rng.integers(256, 1028))
    if flags == 1:
        cols = [1, 5, 33, 300][words.bit_length() - 1]   # tiny and ragged rows
    full = rng.integers(0, 2**32, size=(rows, cols, words), dtype=np.uint64).astype(np.uint32)
    pool = rng.integers(0, 2**32, size=(64, words), dtype=np.uint64).astype(np.uint32)
    pooled = pool[rng.integers(0, 64, size=(rows, cols))]
    # low-entropy descriptors: many equal costs, exercising the tie rules
    sparse = (full & rng.integers(0, 2**32, size=full.shape, dtype=np.uint64).astype(np.uint32)
              & rng.integers(0, 2**32, size=full.shape, dtype=np.uint64).astype(np.uint32))
    for d0, d1 in ((full, full[:, ::-1].copy()), (pooled, pool[rng.integers(0, 64, size=(rows, cols))]),
                   (sparse, sparse[:, rng.permutation(cols)].copy())):
        for lr in (0, 3):
            exp = oracle.search(d0, d1, flags, lr)
            got = host(gpu.search(dev(d0), dev(d1), flags, lr))
            assert np.array_equal(got, exp), lr


def test_search_chunked_rows(gpu, oracle, search_impl):
    """rows wider than one LDS chunk (VALU: 2048 256-bit descriptors; MFMA:
    224 expanded 256-bit descriptors), with a ragged last tile"""
    rng = np.random.default_rng(9)
    rows, cols, words = 3, 4500, 8
    pool = rng.integers(0, 2**32, size=(300, words), dtype=np.uint64).astype(np.uint32)
    d0 = pool[rng.integers(0, 300, size=(rows, cols))]
    d1 = pool[rng.integers(0, 300, size=(rows, cols))]
    for flags in (1, 2, 3):
        assert np.array_equal(host(gpu.search(dev(d0), dev(d1), flags, 2)),
                              oracle.search(d0, d1, flags, 2))


def test_search_fp4_random_sweep(gpu):
    """The fp4 matrix-core search (default for 64- to 256-bit descriptors)
    against the vector-ALU and int8 kernels, which the tests above pin to the
    oracle: 150 seeded cases over row widths around the tile (32) and around
    one and two LDS chunks of either matrix-core kernel (the chunk sizes the
    library reports for each descriptor width) up to 4096, every flag set and
    L/R limit, and random, duplicate-heavy and sparse descriptors."""
    rng = np.random.default_rng(777)
    edges = {}
    for words in (2, 4, 8):
        e = {1, 2, 31, 32, 33, 63, 4095, 4096}
        for impl in (gpu.SEARCH_MFMA, gpu.SEARCH_MFMA_I8):
            ch = gpu.search_chunk_cols(impl, words)
            assert ch > 0
            e |= {ch - 1, ch, ch + 1, 2 * ch - 1, 2 * ch, 2 * ch + 1}
        edges[words] = sorted(w for w in e if 1 <= w <= 4096)
    assert 1920 in edges[4] and 3328 in edges[2]   # the 1920-column row is two fp4 chunks
    try:
        for case in range(150):
            words = int(rng.choice([2, 4, 8]))
            W = int(rng.choice(edges[words])) if rng.random() < 0.6 else int(rng.integers(1, 4097))
            H = int(rng.integers(1, 4))
            flags, lr = int(rng.integers(0, 4)), int(rng.integers(0, 5))
            kind = rng.integers(0, 3)
            if kind == 0:
                d0 = rng.integers(0, 2**32, size=(H, W, words), dtype=np.uint64).astype(np.uint32)
                d1 = rng.integers(0, 2**32, size=(H, W, words), dtype=np.uint64).astype(np.uint32)
            elif kind == 1:
                pool = rng.integers(0, 2**32, size=(max(2, W // 8), words), dtype=np.uint64).astype(np.uint32)
                d0 = pool[rng.integers(0, len(pool), size=(H, W))]
                d1 = pool[rng.integers(0, len(pool), size=(H, W))]
            else:
                m = lambda: rng.integers(0, 2**32, size=(H, W, words), dtype=np.uint64).astype(np.uint32)
                d0, d1 = m() & m() & m(), m() & m() & m()
            q0, q1 = dev(d0), dev(d1)
            got = {}
            for name, impl in (("fp4", gpu.SEARCH_MFMA), ("i8", gpu.SEARCH_MFMA_I8), ("valu", gpu.SEARCH_VALU)):
                gpu.set_search_impl(impl)
                got[name] = host(gpu.search(q0, q1, flags, lr))
            assert np.array_equal(got["fp4"], got["valu"]), (case, W, H, words, flags, lr, kind)
            assert np.array_equal(got["i8"], got["valu"]), (case, W, H, words, flags, lr, kind)
    finally:
        gpu.set_search_impl(gpu.SEARCH_MFMA)


def test_search_planted_duplicates(gpu, oracle, search_impl):
    cols = 700
    rng = np.random.default_rng(5)
    d1 = rng.integers(0, 2**32, size=(1, cols, 4), dtype=np.uint64).astype(np.uint32)
    d0 = d1.copy()
    d1[0, 650] = d1[0, 10]          # left pixel 10 now has two exact matches
    got = host(gpu.search(dev(d0), dev(d1), 1))
    assert got[0, 10] == -32768 and got[0, 11] == 0
    assert np.array_equal(got, oracle.search(d0, d1, 1))
    got0 = host(gpu.search(dev(d0), dev(d1), 0))
    assert got0[0, 10] == 0


# ---------------------------------------------------------------- agree

@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("double", [False, True])
def test_agree_exact(gpu, oracle, dtype, double):
    """test/agree.cu: n = 10, random images and disparities in [-1, W),
    random threshold and minimum variance; int16 output and corrmap exact."""
    import torch
    rng = np.random.default_rng(2000 + double + 2 * (dtype == np.uint16))
    for n in (10, 12, 33, 64, 7):
        rows, cols = 24, int(rng.integers(1024, 2048))
        s0 = rand_stack(rng, n, rows, cols, dtype)
        s1 = rand_stack(rng, n, rows, cols, dtype)
        disp = rng.integers(-1, cols, size=(rows, cols)).astype(np.int16)
        disp[rng.random((rows, cols)) < 0.1] = -32768
        thr = float(rng.uniform(-0.9, 0.9))
        for minvar in (float(rng.uniform(0, 75)), None):
            mv_n = None if minvar is None else np.float32(minvar) * np.float32(n)
            exp, exp_cm = oracle.agree(disp, s0, s1, thr, mv_n, double=double, corrmap=True)
            d = dev(disp)
            cm = torch.empty((rows, cols), dtype=torch.float64 if double else torch.float32,
                             device="cuda")
            gpu.agree(d, dev(s0), dev(s1), thr, minvar, 1 if double else 0, corrmap=cm)
            assert np.array_equal(host(d), exp), (n, minvar)
            assert same(host(cm), exp_cm), (n, minvar)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("double", [False, True])
def test_agree_subpixel_exact(gpu, oracle, dtype, double):
    import torch
    rng = np.random.default_rng(3000 + double + 2 * (dtype == np.uint16))
    for n, step in ((10, 0.25), (33, 0.1), (12, 0.05), (64, 0.05), (5, 0.3)):
        rows, cols = 12, int(rng.integers(1024, 1600))
        # smooth random rows so that interpolation sits near real correlation peaks
        base = rand_stack(rng, n, rows, cols + 8, dtype)
        s0 = base[:, :, 4:cols + 4].copy()
        s1 = base[:, :, 2:cols + 2].copy()
        disp = np.full((rows, cols), 2, np.int16)
        mix = rng.random((rows, cols))
        disp[mix < 0.3] = rng.integers(-1, cols, size=int((mix < 0.3).sum()))
        disp[mix > 0.95] = -32768
        disp[:, 0] = 0                       # col1 == 0 border path
        disp[:, -1] = 0                      # col1 == W-1 border path
        thr = float(rng.uniform(-0.5, 0.9))
        for minvar in (float(rng.uniform(0, 75)), None):
            mv_n = None if minvar is None else np.float32(minvar) * np.float32(n)
            exp, exp_cm = oracle.agree_subpixel(disp, s0, s1, thr, step, mv_n, double=double,
                                                corrmap=True)
            cm = torch.empty((rows, cols), dtype=torch.float64 if double else torch.float32,
                             device="cuda")
            out = gpu.agree_subpixel(dev(disp), dev(s0), dev(s1), thr, step, minvar,
                                     1 if double else 0, corrmap=cm)
            got, got_cm = host(out), host(cm)
            assert same(got, exp), (n, step, minvar)
            assert same(got_cm, exp_cm), (n, step, minvar)


@pytest.fixture
def scan_subpixel(gpu):
    """the full-scan subpixel kernel for one test, then back to the default"""
    gpu.set_subpixel_impl(gpu.SUBPIXEL_SCAN)
    yield
    gpu.set_subpixel_impl(gpu.SUBPIXEL_PRUNED)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("double", [False, True])
def test_agree_subpixel_pruned_exact(gpu, oracle, dtype, double):
    """Precision::SINGLE / DOUBLE without a corrmap run the pruned kernel: bit-exact
    against the oracle on smooth rows (multi-candidate peaks), saturated rows
    (parabolas leaving the sample range: every grid point scored), constant
    and low-variance pixels, thresholds down to -1, and grid sizes 3..64
    (101 points fall back to the full scan)."""
    rng = np.random.default_rng(3100 + (dtype == np.uint16) + 2 * double)
    hi = np.iinfo(dtype).max
    cases = ((10, 0.25, "smooth"), (33, 0.1, "smooth"), (12, 0.05, "smooth"), (64, 0.05, "smooth"),
             (5, 0.3, "smooth"), (33, 0.1, "saturated"), (20, 0.1, "flat"), (33, 0.02, "smooth"),
             (2, 0.5, "smooth"), (40, 0.0317, "smooth"))
    for n, step, kind in cases:
        rows, cols = 12, int(rng.integers(700, 1300))
        x = np.arange(cols + 8, dtype=np.float64)
        base = np.empty((n, rows, cols + 8))
        for t in range(n):
            for r in range(rows):
                ph, fr = rng.uniform(0, 6.3, 3), rng.uniform(0.05, 0.6, 3)
                base[t, r] = sum(np.sin(fr[i] * x + ph[i]) for i in range(3))
        base = (base - base.min()) / (base.max() - base.min())
        if kind == "saturated":
            base = np.clip(base * 1.6 - 0.3, 0, 1)
        base = np.rint(base * hi).astype(dtype)
        if kind == "flat":   # constant and nearly constant pixels
            base[:, :, ::7] = base[0:1, :, ::7]
            base[:, :, 3::7] = (base[0:1, :, 3::7].astype(np.int64)
                                + rng.integers(0, 2, (n, rows, 1))).clip(0, hi)
        shift = 2
        s0 = base[:, :, 4:cols + 4].copy()
        s1 = base[:, :, 4 - shift:cols + 4 - shift].copy()
        # noise so that correlations are not all 1
        noise = rng.normal(0, hi * 0.004, s1.shape)
        s1 = np.clip(np.rint(s1 + noise), 0, hi).astype(dtype)
        disp = np.full((rows, cols), shift, np.int16)
        mix = rng.random((rows, cols))
        disp[mix < 0.2] = rng.integers(-1, cols, size=int((mix < 0.2).sum()))
        disp[mix > 0.95] = -32768
        disp[:, 0] = 0
        disp[:, -1] = 0
        for thr in (float(rng.uniform(0.3, 0.95)), 0.999, -1.0):
            for minvar in (float(rng.uniform(0, 5)), None):
                mv_n = None if minvar is None else np.float32(minvar) * np.float32(n)
                exp, _ = oracle.agree_subpixel(disp, s0, s1, thr, step, mv_n, double=double)
                got = host(gpu.agree_subpixel(dev(disp), dev(s0), dev(s1), thr, step, minvar,
                                              1 if double else 0))
                bad = np.argwhere(~((got == exp) | (np.isnan(got) & np.isnan(exp))))
                assert same(got, exp), (n, step, kind, thr, minvar, bad[:5])


def test_pruned_equals_scan_random_sweep(gpu):
    """300 seeded random configurations (u8/u16, n 2..65, 8 grid steps,
    thresholds -1..0.999, with and without min_variance, FP32 and FP64;
    smooth, uniform, saturated, flat, low-contrast and scene inputs with
    random, invalid and out-of-range disparities): the pruned kernel equals
    the full-scan kernel (itself bit-exact against the oracle) bit for bit."""
    from libbicos_amd import scene
    rng = np.random.default_rng(4242)
    try:
        for case in range(300):
            dtype = np.uint16 if rng.random() < 0.6 else np.uint8
            hi = np.iinfo(dtype).max
            n = int(rng.choice([2, 3, 4, 5, 8, 10, 12, 16, 20, 24, 33, 40, 48, 64, 65]))
            step = float(rng.choice([0.5, 0.3, 0.25, 0.2, 0.1, 0.05, 0.0317, 0.035]))
            rows, cols = int(rng.integers(4, 24)), int(rng.integers(64, 700))
            kind = rng.choice(["smooth", "uniform", "saturated", "flat", "scene", "lowcontrast"])
            if kind == "scene":
                spec = scene.SCENE_1920x1200_U16 if dtype == np.uint16 else scene.SCENE_640x480_U8
                s0, s1, _ = scene.generate(spec, int(rng.integers(1, 1 << 30)), frames=n, height=rows, width=cols)
                disp = rng.integers(0, 40, size=(rows, cols)).astype(np.int16)
            else:
                x = np.arange(cols + 8, dtype=np.float64)
                base = np.zeros((n, rows, cols + 8))
                for t in range(n):
                    ph, fr = rng.uniform(0, 6.3, (rows, 1, 3)), rng.uniform(0.05, 0.8, (rows, 1, 3))
                    base[t] = np.sin(fr * x[None, :, None] + ph).sum(-1)
                if kind == "uniform":
                    base = rng.random(base.shape)
                base = (base - base.min()) / max(base.max() - base.min(), 1e-9)
                if kind == "saturated":
                    base = np.clip(base * 1.8 - 0.4, 0, 1)
                if kind == "lowcontrast":
                    base = 0.5 + 0.002 * (base - 0.5)
                base = np.rint(base * hi).astype(dtype)
                if kind == "flat":
                    base[:, :, ::5] = base[0:1, :, ::5]
                sh = int(rng.integers(0, 4))
                s0 = base[:, :, 4:cols + 4].copy()
                noise = rng.normal(0, hi * float(rng.choice([0, 0.001, 0.01])), s0.shape)
                s1 = np.clip(np.rint(base[:, :, 4 - sh:cols + 4 - sh] + noise), 0, hi).astype(dtype)
                disp = np.full((rows, cols), sh, np.int16)
            mix = rng.random((rows, cols))
            disp[mix < 0.15] = rng.integers(-1, cols, size=int((mix < 0.15).sum()))
            disp[mix > 0.95] = -32768
            thr = float(rng.choice([-1.0, 0.0, 0.5, 0.9, 0.96, 0.99, 0.999]))
            minvar = None if rng.random() < 0.3 else float(rng.choice([0.5, 2.0, 10.0, 100.0]))
            prec = int(rng.integers(0, 2))
            d_d, s0_d, s1_d = dev(disp), dev(s0), dev(s1)
            gpu.set_subpixel_impl(gpu.SUBPIXEL_PRUNED)
            a = host(gpu.agree_subpixel(d_d, s0_d, s1_d, thr, step, minvar, prec))
            gpu.set_subpixel_impl(gpu.SUBPIXEL_SCAN)
            b = host(gpu.agree_subpixel(d_d, s0_d, s1_d, thr, step, minvar, prec))
            assert same(a, b), (case, dtype, n, step, kind, thr, minvar, prec)
    finally:
        gpu.set_subpixel_impl(gpu.SUBPIXEL_PRUNED)


def test_agree_subpixel_scan_knob(gpu, oracle, scan_subpixel):
    """the full-scan kernel stays selectable and agrees with the oracle"""
    rng = np.random.default_rng(3200)
    n, rows, cols = 33, 8, 900
    s0 = rand_stack(rng, n, rows, cols, np.uint16)
    s1 = np.roll(s0, 3, axis=2)
    disp = np.full((rows, cols), 3, np.int16)
    exp, _ = oracle.agree_subpixel(disp, s0, s1, 0.5, 0.1, None)
    got = host(gpu.agree_subpixel(dev(disp), dev(s0), dev(s1), 0.5, 0.1, None, 0))
    assert same(got, exp)


# ----------------------------------------------------------- whole path

def run_match(gpu, L, R, cfg):
    d, cm = gpu.match_device(dev(L), dev(R), cfg, want_corrmap=True)
    return host(d), (host(cm) if cm is not None else None)


def oracle_match(oracle, L, R, cfg, gpu):
    cons = None
    if isinstance(cfg.variant, gpu.Variant.Consistency):
        cons = (cfg.variant.max_lr_diff, cfg.variant.no_dupes)
    return oracle.match(L, R, thr=cfg.nxcorr_threshold, step=cfg.subpixel_step,
                        min_variance=cfg.min_variance, full=cfg.mode == 1,
                        double=cfg.precision == 1, consistency=cons, corrmap=True)


def test_match_640x480_u8_full(gpu, oracle):
    """640x480 u8, 12 frames, FULL, threshold 0.9, no subpixel, NoDuplicates."""
    from libbicos_amd import scene
    L, R, gt = scene.generate(scene.SCENE_640x480_U8, seed=101)
    cfg = gpu.Config(nxcorr_threshold=0.9, mode=gpu.TransformMode.Full)
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert got.dtype == np.int16 and np.array_equal(got, exp)
    assert same(got_cm, exp_cm)
    v = got != -32768
    assert (np.abs(got[v] - gt[v]) <= 1).mean() > 0.95


def test_match_1920_band_u16_limited_subpixel(gpu, oracle):
    """1920 wide u16, 33 frames, LIMITED, min_variance 2, threshold 0.96, step 0.1."""
    from libbicos_amd import scene
    L, R, gt = scene.generate(scene.SCENE_1920x1200_U16, seed=102, height=64)
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert got.dtype == np.float32
    assert same(got, exp) and same(got_cm, exp_cm)
    d, _ = gpu.match_device(dev(L), dev(R), cfg, want_corrmap=False)  # pruned subpixel
    assert same(host(d), exp)
    v = ~np.isnan(got)
    assert v.mean() > 0.5 and (np.abs(got[v] - gt[v]) <= 1).mean() > 0.95


@pytest.mark.parametrize("thr", [0.5, 0.75, 0.9])
@pytest.mark.parametrize("variant", ["nodup", "cons1", "cons2dup"])
def test_match_integration_matrix(gpu, oracle, thr, variant):
    """test/integration.cu:45-87: FULL mode on 16 frames (256-bit descriptors),
    thresholds x variants x subpixel steps {none, 0.1, 0.25, 0.5}."""
    from libbicos_amd import scene
    spec = scene.SceneSpec(720, 48, 16, 8, 5.0, 90.0, 20.0, 235.0, 2.0)
    L, R, _ = scene.generate(spec, seed=103)
    var = {"nodup": gpu.Variant.NoDuplicates(), "cons1": gpu.Variant.Consistency(1, False),
           "cons2dup": gpu.Variant.Consistency(2, True)}[variant]
    for step in (None, 0.1, 0.25, 0.5):
        cfg = gpu.Config(nxcorr_threshold=thr, subpixel_step=step, mode=gpu.TransformMode.Full,
                         variant=var)
        got, got_cm = run_match(gpu, L, R, cfg)
        exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
        assert same(got, exp), step
        assert same(got_cm, exp_cm), step


def test_match_4096_band_u16_64frames_consistency(gpu, oracle, search_impl):
    """4096 wide u16, 64 frames (256-bit), Consistency(max_lr_diff 1), step 0.05."""
    from libbicos_amd import scene
    L, R, gt = scene.generate(scene.SCENE_4096x3000_U16, seed=104, height=6)
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.05, min_variance=2.0,
                     variant=gpu.Variant.Consistency(1, False))
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert same(got, exp) and same(got_cm, exp_cm)
    d, _ = gpu.match_device(dev(L), dev(R), cfg, want_corrmap=False)  # pruned subpixel
    assert same(host(d), exp)
    v = ~np.isnan(got)
    assert v.mean() > 0.5 and (np.abs(got[v] - gt[v]) <= 1).mean() > 0.95


def test_match_precision_double_and_no_threshold(gpu, oracle):
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=105, height=16, width=800, frames=20)
    for cfg in (gpu.Config(nxcorr_threshold=0.8, subpixel_step=0.2, precision=gpu.Precision.Double),
                gpu.Config(nxcorr_threshold=0.8, precision=gpu.Precision.Double, min_variance=5.0),
                gpu.Config(nxcorr_threshold=None),
                gpu.Config(nxcorr_threshold=None, subpixel_step=0.1)):
        got, got_cm = run_match(gpu, L, R, cfg)
        exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
        assert same(got, exp)
        assert (got_cm is None) == (exp_cm is None)
        if got_cm is not None:
            assert same(got_cm, exp_cm)
        d, _ = gpu.match_device(dev(L), dev(R), cfg, want_corrmap=False)  # pruned when subpixel
        assert same(host(d), exp)


def test_match_full_size_1920x1200x33(gpu, oracle):
    """The whole 1920x1200x33 u16 workload, bit-exact against the oracle."""
    from libbicos_amd import scene
    L, R, gt = scene.generate(scene.SCENE_1920x1200_U16, seed=scene.BENCH_SEED)
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert same(got, exp) and same(got_cm, exp_cm)
    # without a corrmap the subpixel stage is the pruned kernel
    d, cm = gpu.match_device(dev(L), dev(R), cfg, want_corrmap=False)
    assert cm is None and same(host(d), exp)
    v = ~np.isnan(got)
    print(f"valid {v.mean():.4f}, within 1px of ground truth {(np.abs(got[v] - gt[v]) <= 1).mean():.4f}")


@pytest.mark.timeout(900)
def test_match_full_size_4096x3000x64_consistency(gpu, oracle):
    """The whole 4096x3000x64 u16 stress workload (256-bit descriptors,
    Consistency(max_lr_diff 1), step 0.05, threshold 0.96, min_variance 2),
    bit-exact against the oracle on every host core (about a minute)."""
    import torch
    from libbicos_amd import scene, window
    spec = scene.SCENE_4096x3000_U16
    dL, dR, _ = window.scene_device(spec, scene.BENCH_SEED)
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.05, min_variance=2.0,
                     variant=gpu.Variant.Consistency(1, False))
    d, _ = gpu.match_device(dL, dR, cfg)
    got = host(d)
    L, R = host(dL), host(dR)
    del dL, dR, d
    torch.cuda.empty_cache()
    print("oracle on the full 4096x3000x64 pair ...", flush=True)
    exp, _ = oracle.match(L, R, thr=0.96, step=0.05, min_variance=2.0, consistency=(1, False))
    assert same(got, exp)
    v = ~np.isnan(got)
    print(f"valid {v.mean():.4f}")
    assert v.mean() > 0.5


def test_workspace_reuse_across_sizes(gpu, oracle):
    """consecutive calls with shrinking and growing sizes reuse the workspace"""
    from libbicos_amd import scene
    cfg = gpu.Config(nxcorr_threshold=0.9, subpixel_step=0.25, mode=gpu.TransformMode.Full)
    for rows, cols in ((40, 600), (8, 200), (60, 900), (8, 200)):
        L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=rows * cols, height=rows, width=cols)
        got, got_cm = run_match(gpu, L, R, cfg)
        exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
        assert same(got, exp) and same(got_cm, exp_cm)


def test_numpy_match_api(gpu, oracle):
    """pybicos-style host API: lists of numpy images in, numpy out."""
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=106, height=32, width=400)
    cfg = gpu.Config(nxcorr_threshold=0.9, mode=gpu.TransformMode.Full)
    d, cm = gpu.match(list(L), list(R), cfg)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert np.array_equal(d, exp) and same(cm, exp_cm)
    d2, cm2 = gpu.match(list(L), list(R), gpu.Config(nxcorr_threshold=None, mode=1))
    assert cm2 is None and d2.dtype == np.int16


def test_cpp_api_cli(gpu, oracle, tmp_path):
    """BICOS::match through the C++ API (bicos_cli) with the reference CLI's
    option mapping: FULL unless --limited; min_variance only when -v is given
    (src/cli.cpp:154-156)."""
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=107, height=20, width=640, frames=24)
    (tmp_path / "l.raw").write_bytes(L.tobytes())
    (tmp_path / "r.raw").write_bytes(R.tobytes())
    cli = os.path.join(ROOT, "libbicos_amd", "bicos_cli")
    out = tmp_path / "d.raw"
    base = [cli, "--left", str(tmp_path / "l.raw"), "--right", str(tmp_path / "r.raw"),
            "-n", "24", "--rows", "20", "--cols", "640", "--depth", "16", "--limited",
            "-t", "0.9", "-s", "0.1", "-m", "2", "--corrmap", str(tmp_path / "c.raw"), "-o", str(out)]
    for extra, minvar in (([], None), (["-v", "1.5"], 1.5)):
        r = subprocess.run(base + extra, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = np.frombuffer(out.read_bytes(), np.float32).reshape(20, 640)
        got_cm = np.frombuffer((tmp_path / "c.raw").read_bytes(), np.float32).reshape(20, 640)
        exp, exp_cm = oracle.match(L, R, thr=0.9, step=0.1, min_variance=minvar, full=False,
                                   consistency=(2, False), corrmap=True)
        assert same(got, exp) and same(got_cm, exp_cm), extra


def test_python_cli_png_stacks(gpu, oracle, tmp_path):
    """python -m libbicos_amd.cli on numbered PNG stacks (two-folder form,
    16-bit): the raw TIFF disparity and correlation map equal the oracle's,
    and the .xyz point cloud holds the oracle's selection of the reprojected
    points (src/cli.cpp:190-250)."""
    import pointcloud as PC
    from libbicos_amd import imageio, scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=111, height=24, width=500, frames=20)
    for side, st in (("l", L), ("r", R)):
        (tmp_path / side).mkdir()
        for t in range(st.shape[0]):
            imageio.write_png(str(tmp_path / side / f"{t}.png"), st[t])
    Q = np.array([[1, 0, 0, -250.0], [0, 1, 0, -12.0], [0, 0, 0, 800.0], [0, 0, 0.05, 0]])
    (tmp_path / "q.yml").write_text("%YAML:1.0\n---\nQ: !!opencv-matrix\n   rows: 4\n   cols: 4\n"
                                    "   dt: d\n   data: [ " + ", ".join(repr(float(v)) for v in Q.ravel())
                                    + " ]\n")
    out = tmp_path / "o" / "disp.png"
    out.parent.mkdir()
    cmd = [sys.executable, "-m", "libbicos_amd.cli", str(tmp_path / "l"), str(tmp_path / "r"),
           "--limited", "-t", "0.9", "-s", "0.1", "--corrmap", "-q", str(tmp_path / "q.yml"),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    # no -v: no min_variance (src/cli.cpp:154-156)
    exp, exp_cm = oracle.match(L, R, thr=0.9, step=0.1, min_variance=None, full=False, corrmap=True)
    got = imageio.read_tiff(str(tmp_path / "o" / "disp.tiff"))
    got_cm = imageio.read_tiff(str(tmp_path / "o" / "disp-corrmap.tiff"))
    assert same(got, exp) and same(got_cm, exp_cm)
    assert imageio.read_png(str(tmp_path / "o" / "disp.png")).shape == (24, 500, 3)
    pts_exp = PC.select(exp, PC.reproject(exp, Q))[0]
    lines = (tmp_path / "o" / "disp.xyz").read_text().splitlines()
    assert len(lines) == len(pts_exp)
    assert lines[:50] == [f"{x:g} {y:g} {z:g}" for x, y, z in pts_exp[:50].tolist()]


def test_async_streams_and_workspace_growth(gpu, oracle):
    """match_device on two non-default streams in turn, the workspace growing
    between calls (larger images on later calls): each call is ordered after
    the previous user of the shared workspace and equals the oracle, and a
    call that does not grow the workspace returns while work queued ahead of
    it on its stream is still running (no host synchronisation; the
    reference creates, synchronises and destroys streams per call,
    src/impl/cuda.cu.in:83-92)."""
    import torch
    from libbicos_amd import scene
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    sizes = [(16, 256), (24, 640), (16, 256), (40, 1024), (40, 1024), (8, 128)]
    grows = [False, True, False, True, False, False]   # call 0 follows a released workspace
    pairs = [scene.generate(scene.SCENE_1920x1200_U16, seed=300 + k, height=h, width=w, frames=33)[:2]
             for k, (h, w) in enumerate(sizes)]
    exp = [oracle.match(L, R, thr=0.96, step=0.1, min_variance=2.0)[0] for L, R in pairs]
    devs = [(dev(L), dev(R)) for L, R in pairs]
    busy = torch.rand(4096, 4096, device="cuda")
    torch.cuda.synchronize()
    gpu.release_workspace()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs, early = [], []
    for k, (dl, dr) in enumerate(devs):
        s = streams[k % 2]
        with torch.cuda.stream(s):
            x = busy
            for _ in range(12):          # tens of ms of matrix products ahead of the call
                x = torch.mm(x, busy)
            d, _ = gpu.match_device(dl, dr, cfg, stream=s)
            early.append(not s.query())
        outs.append((d, x))
    torch.cuda.synchronize()
    for k, ((d, _), e) in enumerate(zip(outs, exp)):
        assert same(host(d), e), k
    assert all(e for e, g in zip(early, grows) if not g), early


def test_errors_on_device(gpu):
    import torch
    s = torch.zeros((17, 4, 8), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="258 bits"):
        gpu.match_device(s, s, gpu.Config(mode=gpu.TransformMode.Full))
    with pytest.raises(gpu.Exception, match="at least two"):
        gpu.match_device(s[:1], s[:1])
    with pytest.raises(ValueError, match="subpixel_step"):
        gpu.match_device(s[:4], s[:4], gpu.Config(subpixel_step=-0.1))


def test_match_separately_allocated_frames(gpu, oracle):
    """frames in separate allocations (as with one GpuMat per image) are
    packed into a strided stack first; results are unchanged"""
    import torch
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=109, height=24, width=700, frames=20)
    cfg = gpu.Config(nxcorr_threshold=0.9, subpixel_step=0.25, min_variance=1.0)
    fl = [dev(L[t]).clone() for t in range(L.shape[0])]
    fr = [dev(R[t]).clone() for t in range(R.shape[0])][::-1][::-1]
    d, cm = gpu.match_device(fl, fr, cfg, want_corrmap=True)
    exp, exp_cm = oracle_match(oracle, L, R, cfg, gpu)
    assert same(host(d), exp) and same(host(cm), exp_cm)
    # transform alone on a shuffled frame order (non-monotonic addresses)
    perm = [3, 0, 7, 1] + list(range(8, 20)) + [2, 4, 5, 6]
    got = host(gpu.transform([fl[i] for i in perm], 0), np.uint32)
    assert np.array_equal(got, oracle.transform(L[perm], full=False))


def precision_stats(got, got_cm, exp, exp_cm):
    fg, fe = np.isfinite(got), np.isfinite(exp)
    both = fg & fe
    dd = np.abs(got[both] - exp[both])
    cmb = np.isfinite(got_cm) & np.isfinite(exp_cm)
    dn = np.abs(got_cm[cmb].astype(np.float64) - exp_cm[cmb])
    flips = (fg != fe).sum() / max(1, (fg | fe).sum())
    return dict(valid_both=int(both.sum()), d_max=float(dd.max()), d_mean=float(dd.mean()),
                d_over_0p1=float((dd > 0.1 + 1e-4).mean()), ncc_max=float(dn.max()),
                ncc_mean=float(dn.mean()), flip_frac=float(flips))


def test_precision_half_mfma_vs_fp64_oracle(gpu, oracle):
    """HALF_MFMA subpixel mode on the 1920-wide 33-frame u16 workload: the
    error of each result against the FP64 oracle (Precision::DOUBLE
    semantics, agree.cuh:35-65) is measured and bounded."""
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=110, height=96)
    cfg = gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0,
                     precision=gpu.Precision.HalfMFMA)
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle.match(L, R, thr=0.96, step=0.1, min_variance=2.0, double=True,
                               corrmap=True)
    st = precision_stats(got, got_cm, exp, exp_cm)
    print("HALF_MFMA vs FP64:", st)
    assert st["valid_both"] > 0.5 * got.size
    # the grid spans 2 px; a flat two-peaked curve may put the approximate
    # argmax on the other peak, which the runner-up rescoring mostly catches
    assert st["d_mean"] < 0.01 and st["d_over_0p1"] < 0.01 and st["d_max"] <= 2.0
    assert st["ncc_mean"] < 1e-3
    assert st["flip_frac"] < 0.005
    # integer stages are shared with the exact path
    plain, _ = run_match(gpu, L, R, gpu.Config(nxcorr_threshold=0.96, subpixel_step=0.1,
                                               min_variance=2.0))
    assert (np.isfinite(plain) | ~np.isfinite(got)).mean() > 0.99


@pytest.mark.parametrize("which", ["1920x1200_u16", "640x480_u8"])
def test_precision_moments_vs_fp64_oracle(gpu, oracle, which):
    """MOMENTS subpixel mode: disparity error of each result against the FP64
    oracle (Precision::DOUBLE semantics, agree.cuh:35-65), measured and
    bounded; with a corrmap asked for, the exact FP32 stage runs and the
    result is bit-identical to it."""
    from libbicos_amd import scene
    if which == "1920x1200_u16":
        L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=112, height=96)
        kw = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
        okw = dict(thr=0.96, step=0.1, min_variance=2.0)
    else:
        L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=113, height=64)
        kw = dict(nxcorr_threshold=0.9, subpixel_step=0.25, mode=gpu.TransformMode.Full)
        okw = dict(thr=0.9, step=0.25, full=True)
    cfg = gpu.Config(precision=gpu.Precision.Moments, **kw)
    got = host(gpu.match_device(dev(L), dev(R), cfg)[0])
    exp, exp_cm = oracle.match(L, R, double=True, corrmap=True, **okw)
    st = precision_stats(got, exp_cm, exp, exp_cm)
    print(f"MOMENTS {which} vs FP64:", st)
    assert st["valid_both"] > 0.5 * got.size
    # one grid step where two grid points are within the rounding bound; the
    # 8-bit samples round coarsely, so their bound is wider
    assert st["d_mean"] < 0.02 and st["d_over_0p1"] < (0.01 if which == "1920x1200_u16" else 0.08)
    assert st["d_max"] <= 2.0
    assert st["flip_frac"] == 0.0
    with_cm, cm = run_match(gpu, L, R, cfg)
    ex, ex_cm = run_match(gpu, L, R, gpu.Config(**kw))
    assert same(with_cm, ex) and same(cm, ex_cm)


@pytest.mark.parametrize("which", ["1920x1200_u16", "640x480_u8"])
def test_precision_moments_decisions_exact(gpu, oracle, which):
    """MOMENTS without a corrmap settles accept/reject by a bound on what the
    rounding can change and rescores only where the bound cannot decide:
    the valid set is exactly the exact FP32 path's, and the grid point
    matches it for nearly every pixel."""
    from libbicos_amd import scene
    if which == "1920x1200_u16":
        L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=114, height=96)
        kw = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
    else:
        L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=115, height=64)
        kw = dict(nxcorr_threshold=0.9, subpixel_step=0.25, mode=gpu.TransformMode.Full)
    d = gpu.match_device(dev(L), dev(R), gpu.Config(precision=gpu.Precision.Moments, **kw))[0]
    got = host(d)
    ex, _ = run_match(gpu, L, R, gpu.Config(**kw))
    assert np.array_equal(np.isfinite(got), np.isfinite(ex))
    v = np.isfinite(ex)
    same = (got[v] == ex[v]).mean()
    print(f"MOMENTS {which} without corrmap: same grid point for {same:.4f} of valid pixels, "
          f"max |d| {np.abs(got[v] - ex[v]).max():.2f}")
    assert same > (0.99 if which == "1920x1200_u16" else 0.9)


@pytest.mark.parametrize("which", ["1920x1200_u16", "640x480_u8"])
def test_precision_moments_fast_vs_exact(gpu, oracle, which):
    """MOMENTS_FAST (closed form decided on its own values) against the exact
    FP32 path on a 1920-wide crop: the share of pixels whose accept/reject
    decision flips is at most 0.1 % (u16) / 2 % (u8, one grid step of 0.25
    px), no valid pixel moves by more than one grid step for 99.5 % of the
    pixels valid in both, and the result does not depend on the corrmap
    request being absent (with a corrmap the exact stage runs)."""
    from libbicos_amd import scene
    if which == "1920x1200_u16":
        L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=116, height=96)
        kw = dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0)
        flip_max, step = 1e-3, 0.1
    else:
        L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=117, height=64)
        kw = dict(nxcorr_threshold=0.9, subpixel_step=0.25, mode=gpu.TransformMode.Full)
        flip_max, step = 2e-2, 0.25
    got = host(gpu.match_device(dev(L), dev(R), gpu.Config(precision=gpu.Precision.MomentsFast, **kw))[0])
    ex, _ = run_match(gpu, L, R, gpu.Config(**kw))
    flips = (np.isfinite(got) != np.isfinite(ex)).mean()
    both = np.isfinite(got) & np.isfinite(ex)
    near = (np.abs(got[both] - ex[both]) <= step * 1.0001).mean()
    print(f"MOMENTS_FAST {which}: decision flips {flips:.2e}, within one grid step {near:.5f}, "
          f"exact grid point {(got[both] == ex[both]).mean():.4f}")
    assert flips <= flip_max and near >= 0.995
    # with a corrmap: the exact FP32 stage (equal to SINGLE bit for bit)
    d_cm, cm = gpu.match_device(dev(L), dev(R), gpu.Config(precision=gpu.Precision.MomentsFast, **kw),
                                want_corrmap=True)
    assert same(host(d_cm), ex)


def closed_form_decisions(oracle, L, R, raw, thr, step, minvar_n, margin):
    """What MOMENTS_FAST computes, in double precision from the oracle's
    closed form: (expected float disparity, robust mask). A pixel is robust
    where its decision cannot turn on the kernel's FP32 evaluation of the
    closed form: best score farther than margin from the threshold and from
    every other grid point's score, variances farther than 1 % from minvar."""
    n, rows, cols = L.shape
    grid = oracle.subpixel_grid(step)
    exp = np.full((rows, cols), np.nan, np.float32)
    robust = np.zeros((rows, cols), bool)
    rr, cc = np.nonzero(raw != -32768)
    c1 = cc - raw[rr, cc].astype(np.int64)
    keep = (0 <= c1) & (c1 < cols)
    rr, cc, c1 = rr[keep], cc[keep], c1[keep]
    left = L[:, rr, cc].T
    ys = np.stack([R[:, rr, np.clip(c1 + j, 0, cols - 1)].T for j in (-1, 0, 1)], axis=1)
    nxc, var1 = oracle.subpixel_closed_form(left, ys, step)
    lf = left.astype(np.float64)
    var0 = ((lf - lf.mean(1, keepdims=True)) ** 2).sum(1)
    border = (c1 == 0) | (c1 == cols - 1)
    # border columns: the plain correlation at col1, which is the closed form at x = 0
    k0 = int(np.argmin(np.abs(grid)))
    scores = np.where(var1 < minvar_n, -1.0, nxc) if minvar_n is not None else nxc
    # 0 / 0 never beats the best; nor a right variance that is zero up to
    # rounding (below 2^-30 of the three columns' total variance)
    total = var1[:, 0] + var1[:, int(np.argmin(np.abs(grid)))] + var1[:, -1]
    scores = np.where(np.isnan(scores) | (var1 <= 2.0 ** -30 * total[:, None]), -np.inf, scores)
    low0 = (var0 < minvar_n) if minvar_n is not None else np.zeros(len(var0), bool)
    best_k = np.argmax(scores, axis=1)                      # first maximum
    best = scores[np.arange(len(best_k)), best_k]
    srt = np.sort(scores, axis=1)
    gap = np.where(best_k >= 0, srt[:, -1] - srt[:, -2], 1.0)
    best = np.where(low0, -1.0, best)
    x = np.where(low0 | border, np.float32(0), grid[best_k])
    best = np.where(border & ~low0, scores[:, k0], best)
    accept = best >= thr
    d = raw[rr, cc].astype(np.float32)
    exp[rr, cc] = np.where(accept, d - x, np.nan)
    rob = np.abs(best - thr) > margin
    rob &= border | low0 | (gap > margin)
    if minvar_n is not None:
        rob &= np.abs(var0 - minvar_n) > 0.01 * minvar_n
        rob &= (np.abs(var1 - minvar_n) > 0.01 * minvar_n).all(1)
    robust[rr, cc] = rob
    # pixels without a valid match are NaN on every path
    robust[raw == -32768] = True
    return exp, robust


@pytest.mark.parametrize("which", ["1920x1200_u16", "640x480_u8"])
def test_moments_fast_follows_closed_form(gpu, oracle, which):
    """The MOMENTS_FAST kernel (moments summed on the GPU, the parabola's
    coefficients and the quartic variance evaluated in FP32) decides as the
    closed form evaluated in double from its definition by the oracle, given
    the oracle's integer search: on every pixel whose decision has a margin
    above the FP32 evaluation error, bit for bit; a wrong coefficient of the
    closed form moves the best grid point or the decision of most pixels."""
    from libbicos_amd import scene
    if which == "1920x1200_u16":
        L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=118, height=48)
        thr, step, mv, full = 0.96, 0.1, 2.0, False
    else:
        L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=119, height=48)
        thr, step, mv, full = 0.9, 0.25, None, True
    n = L.shape[0]
    raw = oracle.search(oracle.transform(L, full), oracle.transform(R, full), 1)
    # margin 1e-5: far above the kernel's evaluation error (the closed form
    # in FP64 from exact integer coefficients, then two conversions to FP32,
    # rsq and two products: a few 1e-7 relative)
    exp, robust = closed_form_decisions(oracle, L, R, raw, thr, step, None if mv is None else mv * n, 1e-5)
    cfg = gpu.Config(nxcorr_threshold=thr, subpixel_step=step, min_variance=mv,
                     mode=gpu.TransformMode.Full if full else gpu.TransformMode.Limited,
                     precision=gpu.Precision.MomentsFast)
    got = host(gpu.match_device(dev(L), dev(R), cfg)[0])
    valid = np.isfinite(exp)
    print(f"MOMENTS_FAST {which}: {robust.mean():.4f} robust, {valid.mean():.3f} valid")
    assert robust.mean() > 0.9 and (valid & robust).sum() > 0.9 * valid.sum()
    bad = robust & ~(np.isnan(got) & np.isnan(exp) | (got == exp))
    assert not bad.any(), (int(bad.sum()), np.argwhere(bad)[:5].tolist())


def test_precision_half_mfma_u8_full_mode(gpu, oracle):
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_640x480_U8, seed=111, height=64)
    cfg = gpu.Config(nxcorr_threshold=0.9, subpixel_step=0.25, mode=gpu.TransformMode.Full,
                     precision=gpu.Precision.HalfMFMA)
    got, got_cm = run_match(gpu, L, R, cfg)
    exp, exp_cm = oracle.match(L, R, thr=0.9, step=0.25, full=True, double=True, corrmap=True)
    st = precision_stats(got, got_cm, exp, exp_cm)
    print("HALF_MFMA u8 vs FP64:", st)
    assert st["valid_both"] > 0.5 * got.size
    assert st["d_mean"] < 0.05 and st["flip_frac"] < 0.02


# ------------------------------------------------------------ point clouds

@pytest.mark.parametrize("kind", ["int16", "float32"])
def test_pointcloud_matches_oracle(gpu, kind):
    """bicos_hip_reproject / bicos_hip_pointcloud against oracle/pointcloud.py
    (cv::reprojectImageTo3D as src/cli.cpp:239 calls it, then save_pointcloud's
    selection): points bit for bit, the same selection and counts. The Q of a
    rectified rig with a principal-point offset makes W = 0 (non-finite
    points) at d = 40 and z < 0 below it."""
    import pointcloud as PC
    rng = np.random.default_rng(21 + (kind == "float32"))
    rows, cols = 37, 301
    if kind == "int16":
        d = rng.integers(-60, 300, size=(rows, cols)).astype(np.int16)
        d[rng.random((rows, cols)) < 0.1] = -32768
        d[rng.random((rows, cols)) < 0.03] = 40
    else:
        d = rng.uniform(-60.0, 300.0, size=(rows, cols)).astype(np.float32)
        d[rng.random((rows, cols)) < 0.1] = np.nan
        d[rng.random((rows, cols)) < 0.03] = 40.0
    f, cx, cy, B = 812.5, 150.25, 18.5, 0.095
    q32 = 1 / B
    Q = [[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 0, f], [0, 0, q32, -(q32 * 40.0)]]   # W = 0 at d = 40
    exp_pts = PC.reproject(d, Q)
    got_pts = host(gpu.reproject(dev(d), Q))
    assert np.array_equal(got_pts, exp_pts, equal_nan=True)
    for allow in (False, True):
        xyz, c = gpu.pointcloud(dev(d), Q, allow_negative_z=allow)
        exp, n, nonfin, negz = PC.select(d, exp_pts, allow)
        assert (c["written"], c["nonfinite"], c["negative_z"]) == (n, nonfin, negz)
        assert nonfin > 0 and (allow or negz > 0)
        assert np.array_equal(host(xyz), exp)
