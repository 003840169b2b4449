"""Pins the CPU oracle (oracle/bicos_oracle.cpp) before it is trusted as the
checker of the HIP kernels.

The reference ships no golden vectors for this path (its tests are
randomised CUDA-vs-CPU comparisons, test/*.cu, and its regression TIFFs are
missing), and it cannot be compiled here (it needs OpenCV). Each oracle
function is therefore pinned by facts fixed independently of its code:
hand-derived known answers, brute-force re-formulations, closed forms and
invariants, chosen so that a wrong emission order, a wrong comparison, a
dropped term or a flipped sign fails at least one check.
"""
import numpy as np
import pytest


def px(values, dtype=np.uint8):
    """a 1x1-pixel stack with the given temporal samples"""
    return np.array(values, dtype).reshape(-1, 1, 1)


def desc_int(words):
    return sum(int(w) << (32 * k) for k, w in enumerate(np.asarray(words).ravel()))


# ------------------------------------------------- descriptor transform
# Known answers derived by hand from the emission rules of
# include/impl/cpu/descriptor_transform.hpp:31-123 (bit k = k-th comparison).
LIMITED_KNOWN = [
    ([1, 2], 0b1011),                  # a<b, a<av, !(b<av), sentinel -1 < 3
    ([3, 1, 2], 0b1011000),            # t=0 all false; closing bits 1,1,0,1
    ([1, 2, 3], 79),                   # av = 2 exactly: 2 < 2 is false
    ([1, 3, 2, 4], 727),               # first pair-sum bit: S0=4 < S2=6
    ([5, 1, 4, 2, 3], 3128),           # strict <: S0=6 vs 6, S1=5 vs 5 give 0
]
FULL_KNOWN = [
    ([2, 0, 1], 24),                   # no |i-t|>1 pair exists for n = 3
    ([1, 4, 2, 3], 199),               # S0 = S2 = 5 -> both pair bits 0
    ([3, 1, 4, 1, 5], 30378),          # 18 bits incl. 6 pair-sum comparisons
]


@pytest.mark.parametrize("values,expected", LIMITED_KNOWN)
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_limited_known_answers(oracle, values, expected, dtype):
    assert desc_int(oracle.transform(px(values, dtype), full=False)) == expected


@pytest.mark.parametrize("values,expected", FULL_KNOWN)
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_full_known_answers(oracle, values, expected, dtype):
    assert desc_int(oracle.transform(px(values, dtype), full=True)) == expected


def _random_stack(rng, n, rows, cols, dtype, levels=None):
    hi = levels if levels else np.iinfo(dtype).max + 1
    return rng.integers(0, hi, size=(n, rows, cols)).astype(dtype)


@pytest.mark.parametrize("full", [False, True])
def test_bit_budget_and_usage(oracle, full):
    """Exactly required_bits positions are ever emitted (4n-6 limited, n^2-2n+3
    full) and each of them is set by some input."""
    rng = np.random.default_rng(1)
    for n in ([2, 3, 4, 5, 9, 10, 17, 33, 34, 65] if not full else [2, 3, 4, 6, 8, 12, 16]):
        emitted = (4 * n - 6 if n >= 4 else {2: 4, 3: 7}[n]) if not full else n * n - 2 * n + 3
        s = _random_stack(rng, n, 8, 512, np.uint16, levels=4)
        d = oracle.transform(s, full=full)
        seen = 0
        for w in range(d.shape[-1]):
            seen |= int(np.bitwise_or.reduce(d[..., w].ravel())) << (32 * w)
        assert seen == (1 << emitted) - 1, n


def test_constant_pixel(oracle):
    """Constant samples: every comparison is false except the always-true
    sentinel bit of n < 4 (prev pair sum -1 < a+b)."""
    for n, exp in [(2, 1 << 3), (3, 1 << 6), (4, 0), (9, 0), (33, 0)]:
        assert desc_int(oracle.transform(px([77] * n, np.uint16), full=False)) == exp
    for n in [2, 3, 5, 12, 16]:
        assert desc_int(oracle.transform(px([77] * n), full=True)) == 0


@pytest.mark.parametrize("full,n", [(False, 9), (False, 17), (False, 33), (False, 65),
                                    (True, 6), (True, 12), (True, 16)])
def test_u8_to_u16_scaling_invariance(oracle, full, n):
    """Scaling every sample by 256 scales sums and the mean exactly, so the
    descriptor cannot change."""
    rng = np.random.default_rng(n)
    s8 = _random_stack(rng, n, 16, 64, np.uint8)
    a = oracle.transform(s8, full=full)
    b = oracle.transform(s8.astype(np.uint16) * 256, full=full)
    assert np.array_equal(a, b)


def test_descriptor_width_independence(oracle):
    """The same stack in a wider descriptor type gives the same bits, zero above."""
    rng = np.random.default_rng(7)
    s = _random_stack(rng, 9, 8, 40, np.uint16)
    base = oracle.transform(s, full=False, words=1)
    for words in (2, 4, 8):
        d = oracle.transform(s, full=False, words=words)
        assert np.array_equal(d[..., 0], base[..., 0]) and not d[..., 1:].any()


def test_mean_bit(oracle):
    """Bit 2 of the limited descriptor is p[0] < mean: checked against numpy."""
    rng = np.random.default_rng(3)
    s = _random_stack(rng, 33, 16, 64, np.uint16)
    d = oracle.transform(s, full=False)
    av = (s.astype(np.float32).sum(0, dtype=np.float32) / np.float32(33)).astype(np.float32)
    assert np.array_equal(((d[..., 0] >> 2) & 1).astype(bool), s[0] < av)


# ------------------------------------------------------------- search

def brute_search(d0, d1, flags, max_lr):
    """Argmin re-formulation of bicos.hpp:50-113 with numpy: first minimum via
    argmin, duplicates via a count of the minimum cost."""
    rows, cols, _ = d0.shape
    out = np.full((rows, cols), -32768, np.int16)
    for r in range(rows):
        cost = np.bitwise_count(d0[r][:, None, :] ^ d1[r][None, :, :]).sum(-1).astype(np.int64)
        best = cost.argmin(1)
        dup = (cost == cost.min(1, keepdims=True)).sum(1) > 1
        rbest = cost.argmin(0)
        rdup = (cost == cost.min(0, keepdims=True)).sum(0) > 1
        for c in range(cols):
            if flags & 1 and dup[c]:
                continue
            b = best[c]
            if flags & 2:
                if flags & 1 and rdup[b]:
                    continue
                rc = rbest[b]
                if abs(c - rc) > max_lr:
                    continue
                out[r, c] = (c + rc) // 2 - b
            else:
                out[r, c] = c - b
    return out


@pytest.mark.parametrize("words", [1, 2, 4, 8])
@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_search_matches_bruteforce(oracle, words, flags):
    rng = np.random.default_rng(10 * words + flags)
    rows, cols = 6, 150
    # few distinct descriptors so ties and duplicates are frequent
    pool = rng.integers(0, 2**32, size=(40, words), dtype=np.uint64).astype(np.uint32)
    d0 = pool[rng.integers(0, 40, size=(rows, cols))]
    d1 = pool[rng.integers(0, 40, size=(rows, cols))]
    for lr in (0, 3):
        exp = brute_search(d0, d1, flags, lr)
        got = oracle.search(d0, d1, flags, lr)
        assert np.array_equal(got, exp)


def test_search_planted(oracle):
    """Hand-built row: the unique exact match wins; a planted duplicate minimum
    invalidates the pixel under no-duplicates but not without it."""
    cols = 8
    d1 = np.array([[[0xF0F0F0F0], [0x0], [0xFFFF0000], [0x12345678], [0x0000FFFF], [0x1], [0x3], [0x7]]],
                  np.uint32)
    d0 = d1.copy()
    d0[0, 5, 0] = 0x12345678      # matches col 3 exactly -> disparity 2
    d0[0, 6, 0] = 0x2             # cost 1 to col 1 (0x0) and to col 6 (0x3): duplicate minimum
    got1 = oracle.search(d0, d1, 1)
    got0 = oracle.search(d0, d1, 0)
    assert got1[0, 5] == 2 and got0[0, 5] == 2
    assert got1[0, 6] == -32768
    assert got0[0, 6] == 6 - 1      # first minimum
    assert got1[0, 0] == 0 and got1[0, 1] == 0


# ------------------------------------------------------ windowed search
# The learned-window path's search: bicos.hpp:50-113 with the scan limited
# to [k - hw, k + hw], k = col - rint(centre of the pixel's 4 x 4 cell),
# clipped to the row; the Consistency back search over [col - hw, col + hw].

def brute_search_window(d0, d1, flags, max_lr, center, hw):
    """Per-pixel numpy re-formulation: the window's costs, first minimum via
    argmin, duplicates via a count of the minimum."""
    rows, cols, _ = d0.shape
    Hq, Wq = rows // 4, cols // 4
    out = np.full((rows, cols), -32768, np.int16)

    def best_in(q, row, lo, hi):
        if lo > hi:
            return None
        cost = np.bitwise_count(row[lo:hi + 1] ^ q[None, :]).sum(-1).astype(np.int64)
        m = cost.min()
        if flags & 1 and (cost == m).sum() > 1:
            return None
        return lo + int(cost.argmin())

    for r in range(rows):
        for c in range(cols):
            cell = (min(r // 4, Hq - 1), min(c // 4, Wq - 1))
            k = c - int(np.rint(np.float32(center[cell])))
            h = int(hw[cell])
            b = best_in(d0[r, c], d1[r], max(k - h, 0), min(k + h, cols - 1))
            if b is None:
                continue
            if flags & 2:
                rc = best_in(d1[r, b], d0[r], max(c - h, 0), min(c + h, cols - 1))
                if rc is None or abs(c - rc) > max_lr:
                    continue
                out[r, c] = (c + rc) // 2 - b
            else:
                out[r, c] = c - b
    return out


def random_windows(rng, rows, cols, widths=(0, 6, 12)):
    """Cell windows that cover the interior, both row ends (clipped), fully
    outside the row (empty) and exact half-way centres (rint: half to even)."""
    Hq, Wq = rows // 4, cols // 4
    center = rng.uniform(-20, 40, size=(Hq, Wq)).astype(np.float32)
    kind = rng.integers(0, 5, size=(Hq, Wq))
    center[kind == 1] = -cols - 30.0                          # window past the right end
    center[kind == 2] = cols + 30.0                           # window before column 0
    center[kind == 3] = np.round(center[kind == 3]) + 0.5     # ties
    hw = rng.choice(np.array(widths, np.int16), size=(Hq, Wq)).astype(np.int16)
    return center, hw


@pytest.mark.parametrize("words", [1, 2, 4, 8])
@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_search_window_matches_bruteforce(oracle, words, flags):
    rng = np.random.default_rng(700 + 10 * words + flags)
    rows, cols = 9, 70       # a partial last cell row and column (9 = 2*4 + 1, 70 = 17*4 + 2)
    pool = rng.integers(0, 2**32, size=(12, words), dtype=np.uint64).astype(np.uint32)
    d0 = pool[rng.integers(0, 12, size=(rows, cols))]
    d1 = pool[rng.integers(0, 12, size=(rows, cols))]
    center, hw = random_windows(rng, rows, cols)
    for lr in (0, 2):
        exp = brute_search_window(d0, d1, flags, lr, center, hw)
        got = oracle.search_window(d0, d1, flags, lr, center, hw)
        assert np.array_equal(got, exp)
    # the cases the windows were built for occur
    assert (exp == -32768).any() and (exp != -32768).any()


def test_search_window_rint_half_even(oracle):
    """Centres 2.5 and 3.5 round to 2 and 4 (half to even): a single exact
    match at column col - 2 (resp. col - 4) is found with hw = 0."""
    rows, cols = 4, 16
    rng = np.random.default_rng(3)
    d1 = rng.integers(0, 2**32, size=(rows, cols, 2), dtype=np.uint64).astype(np.uint32)
    for c, shift in ((2.5, 2), (3.5, 4), (-2.5, -2), (2.4999, 2), (2.5001, 3)):
        d0 = np.roll(d1, shift, axis=1)
        center = np.full((1, 4), c, np.float32)
        hw = np.zeros((1, 4), np.int16)
        got = oracle.search_window(d0, d1, 1, 0, center, hw)
        cols_ok = np.arange(cols)[(np.arange(cols) - shift >= 0) & (np.arange(cols) - shift < cols)]
        assert (got[:, cols_ok] == shift).all(), (c, shift)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_search_window_full_row_is_plain_search(oracle, flags):
    """A window that spans the row at every pixel (centre 0, hw >= cols) is the
    plain search; so is one centred on each pixel's own column range."""
    rng = np.random.default_rng(40 + flags)
    rows, cols = 8, 96
    pool = rng.integers(0, 2**32, size=(30, 4), dtype=np.uint64).astype(np.uint32)
    d0 = pool[rng.integers(0, 30, size=(rows, cols))]
    d1 = pool[rng.integers(0, 30, size=(rows, cols))]
    center = np.zeros((rows // 4, cols // 4), np.float32)
    hw = np.full(center.shape, 2 * cols, np.int16)
    for lr in (0, 3):
        assert np.array_equal(oracle.search_window(d0, d1, flags, lr, center, hw),
                              oracle.search(d0, d1, flags, lr))


def test_match_windowed_fallback_semantics(oracle):
    """The learned-window match on a scene: with whole-row windows it is the
    plain match for every fallback mode; with useless windows (far off the
    true disparity) fallback ALL gives the plain output at every pixel, NONE
    gives the windowed search followed by the NCC stage, and the fallback set
    is the NCC-rejected pixels minus those whose left variance rejects every
    match."""
    from libbicos_amd import scene
    L, R, _ = scene.generate(scene.SCENE_1920x1200_U16, seed=103, height=24, width=256, frames=12)
    kw = dict(thr=0.9, step=0.25, min_variance=2.0)
    plain, plain_cm = oracle.match(L, R, corrmap=True, **kw)
    Hq, Wq = 6, 64
    whole = (np.zeros((Hq, Wq), np.float32), np.full((Hq, Wq), 600, np.int16))
    for fb in (0, 1, 2):
        got, cm, nfb = oracle.match_windowed(L, R, *whole, fallback=fb, corrmap=True, **kw)
        assert np.array_equal(got, plain, equal_nan=True) and np.array_equal(cm, plain_cm, equal_nan=True)
    off = (np.full((Hq, Wq), -150.0, np.float32), np.full((Hq, Wq), 6, np.int16))
    got_all, _, nfb = oracle.match_windowed(L, R, *off, fallback=1, **kw)
    assert np.array_equal(got_all, plain, equal_nan=True)
    got_none, _, nfb0 = oracle.match_windowed(L, R, *off, fallback=0, **kw)
    words = oracle.descriptor_words(12, False)
    ws = oracle.search_window(oracle.transform(L, False, words), oracle.transform(R, False, words),
                              1, -1, *off)
    exp_none, _ = oracle.agree_subpixel(ws, L, R, 0.9, 0.25, minvar=2.0 * 12)
    assert np.array_equal(got_none, exp_none, equal_nan=True) and nfb0 == 0
    # fallback set = invalid after the windowed pass, minus the hopeless
    # pixels (float64 left variance; no pixel of this scene is within FP32
    # rounding of the limit)
    v0 = L.astype(np.float64).var(axis=0) * 12
    hopeless = v0 < 2.0 * 12
    assert nfb == int((np.isnan(exp_none) & ~hopeless).sum()) > 0
    # UNSURE: only cells with the wide window are searched again
    hw_mixed = np.where(np.arange(Wq)[None, :] % 2 == 0, 6, 12).astype(np.int16) * np.ones((Hq, 1), np.int16)
    got_u, _, nfb_u = oracle.match_windowed(L, R, off[0], hw_mixed, hw_lo=6, fallback=2, **kw)
    got_n, _, _ = oracle.match_windowed(L, R, off[0], hw_mixed, hw_lo=6, fallback=0, **kw)
    wide = np.repeat(np.repeat(hw_mixed == 12, 4, 0), 4, 1)
    assert np.array_equal(got_u[~wide], got_n[~wide], equal_nan=True)
    redo = wide & np.isnan(got_n)
    assert np.array_equal(got_u[redo], plain[redo], equal_nan=True)


# ---------------------------------------------------------------- NCC

def ncc64(a, b, minvar=None):
    a = a.astype(np.float64); b = b.astype(np.float64)
    da, db = a - a.mean(), b - b.mean()
    va, vb = (da * da).sum(), (db * db).sum()
    if minvar is not None and (va < minvar or vb < minvar):
        return -1.0
    return (da * db).sum() / np.sqrt(va * vb)


def test_nxcorr_against_float64(oracle):
    rng = np.random.default_rng(5)
    worst = 0.0
    for n in (2, 5, 10, 33, 64):
        for _ in range(200):
            a = rng.integers(0, 65536, n).astype(np.uint16)
            b = rng.integers(0, 65536, n).astype(np.uint16)
            worst = max(worst, abs(oracle.nxcorr(a, b) - ncc64(a, b)))
            assert abs(oracle.nxcorrd(a, b) - ncc64(a, b)) < 1e-12
    assert worst < 1e-5


def test_nxcorr_closed_forms(oracle):
    a = np.array([10, 200, 35, 4000, 17, 900, 65000, 3], np.uint16)
    assert oracle.nxcorr(a, a) == 1.0                                # identical samples
    b = (a.astype(np.int64) * 1 + 7).astype(np.uint16)
    assert abs(oracle.nxcorr(a, b) - 1.0) < 1e-6                     # affine, positive slope
    c = (65535 - a.astype(np.int64)).astype(np.uint16)
    assert abs(oracle.nxcorr(a, c) + 1.0) < 1e-6                     # anti-correlated
    k = np.full(8, 123, np.uint16)
    assert np.isnan(oracle.nxcorr(a, k))                             # zero variance -> NaN
    assert oracle.nxcorr(a, k, minvar=1.0) == -1.0                   # variance filter
    assert oracle.nxcorr(a, a, minvar=1e30) == -1.0


# ------------------------------------------------------- agree / subpixel

def test_subpixel_grid(oracle):
    """x = -1; x <= 1; x += step accumulated in float32 (agree.hpp:162)."""
    def grid(step):
        xs, x = [], np.float32(-1)
        while x <= np.float32(1):
            xs.append(x)
            x = np.float32(x + np.float32(step))
        return np.array(xs, np.float32)
    for step, count, last in [(0.1, 20, 0.90000015), (0.05, 40, 0.9500003), (0.25, 9, 1.0)]:
        g = oracle.subpixel_grid(step)
        assert len(g) == count and g[0] == -1.0
        assert np.array_equal(g, grid(step))
        assert np.float32(g[-1]) == np.float32(last)


def _linear_right_stack(rng, n, cols):
    alpha = rng.integers(1000, 2000, n)
    beta = 2 * rng.integers(1, 21, n)                # even slopes
    c = np.arange(cols)
    return alpha, beta, (alpha[:, None] + beta[:, None] * c[None, :]).astype(np.uint16)


def test_subpixel_cast_wraps(oracle):
    """The float -> u8/u16 cast of an interpolated sample (agree.hpp:164) is
    undefined in C++ when the parabola leaves the sample range; the oracle
    takes x86-64 GCC's conversion to int32 after rint, keeping the low bits.
    Known answers, hand-computed at x = 0.6 (a saturating cast would give
    65535 / 0 / 255 / 0, a truncating one 4320 / ...), and rint's half to
    even inside the range."""
    # (0, 60000, 65535): -27232.5 * 0.36 + 32767.5 * 0.6 + 60000 = 69856.8 -> 69857 -> 69857 - 65536
    assert oracle.interp_sample(0, 60000, 65535, 0.6, np.uint16) == 4321
    # (65535, 5000, 0): 27767.5 * 0.36 - 19660.5 + 5000 = -4664.2 -> -4664 -> 65536 - 4664
    assert oracle.interp_sample(65535, 5000, 0, 0.6, np.uint16) == 60872
    # (0, 230, 255): -102.5 * 0.36 + 76.5 + 230 = 269.6 -> 270 -> 270 - 256
    assert oracle.interp_sample(0, 230, 255, 0.6, np.uint8) == 14
    # (255, 20, 0): 107.5 * 0.36 - 76.5 + 20 = -17.8 -> -18 -> 256 - 18
    assert oracle.interp_sample(255, 20, 0, 0.6, np.uint8) == 238
    # a line through (1, 2, 3) at x = 0.5 is 2.5 -> 2; (3, 4, 5) at +-0.5 is 4.5 -> 4 and 3.5 -> 4
    assert oracle.interp_sample(1, 2, 3, 0.5) == 2
    assert oracle.interp_sample(3, 4, 5, 0.5) == 4
    assert oracle.interp_sample(3, 4, 5, -0.5) == 4
    # the grid's own end points reproduce the columns
    assert oracle.interp_sample(17, 400, 9000, -1.0) == 17 and oracle.interp_sample(17, 400, 9000, 0.0) == 400


def test_subpixel_half_pixel_shift(oracle):
    """Right rows linear in the column; the left pixel sits half a pixel right
    of column col1: the parabola through col1-1..col1+1 reproduces it exactly at
    x = +0.5, so the refined disparity is d - 0.5 and the correlation is 1.
    A flipped b term would pick x = -0.5, a dropped c term breaks equality."""
    rng = np.random.default_rng(11)
    n, cols = 8, 16
    alpha, beta, R = _linear_right_stack(rng, n, cols)
    L = np.full((n, cols), 1500, np.uint16)
    col, d = 10, 4
    L[:, col] = alpha + beta * (col - d) + beta // 2
    disp = np.full((1, cols), -32768, np.int16)
    disp[0, col] = d
    out, cm = oracle.agree_subpixel(disp, L[:, None, :], R[:, None, :], thr=0.5, step_x=0.25,
                                    corrmap=True)
    assert out[0, col] == np.float32(d - 0.5)
    assert cm[0, col] == 1.0
    assert np.isnan(out[0, :col]).all() and np.isnan(cm[0, :col]).all()


def test_subpixel_integer_shift_and_borders(oracle):
    """An exact integer shift peaks at the grid point closest to 0; matches
    landing on the first/last column take the plain-NCC path and keep d."""
    rng = np.random.default_rng(12)
    n, cols, d = 12, 40, 3
    R = rng.integers(0, 65536, size=(n, 1, cols)).astype(np.uint16)
    L = np.zeros_like(R)
    L[:, :, d:] = R[:, :, :cols - d]
    disp = np.full((1, cols), -32768, np.int16)
    disp[0, d:] = d
    out, cm = oracle.agree_subpixel(disp, L, R, thr=0.9, step_x=0.1, corrmap=True)
    x0 = oracle.subpixel_grid(0.1)[10]                 # 7.45e-08, the point nearest 0
    for c in range(d + 1, cols - 1):
        assert out[0, c] == np.float32(np.float32(d) - x0) and cm[0, c] == 1.0
    assert out[0, d] == d and cm[0, d] == 1.0          # col1 == 0: border path
    # out-of-range match: col - d < 0 -> left invalid, corrmap untouched (NaN)
    disp2 = disp.copy(); disp2[0, 1] = 5
    out2, cm2 = oracle.agree_subpixel(disp2, L, R, thr=0.9, step_x=0.1, corrmap=True)
    assert np.isnan(out2[0, 1]) and np.isnan(cm2[0, 1])


def test_agree_semantics(oracle):
    rng = np.random.default_rng(13)
    n, cols = 10, 32
    R = rng.integers(0, 256, size=(n, 1, cols)).astype(np.uint8)
    L = np.zeros_like(R); L[:, :, 2:] = R[:, :, :-2]
    L[:, :, 20] = 99; R[:, :, 18] = 99                 # constant pair: NaN NCC passes
    disp = np.full((1, cols), 2, np.int16)
    disp[0, 0] = 5                                     # out of range -> invalidated
    disp[0, 1] = -32768
    out, cm = oracle.agree(disp, L, R, thr=0.95, corrmap=True)
    assert out[0, 0] == -32768 and np.isnan(cm[0, 0])
    assert out[0, 1] == -32768 and np.isnan(cm[0, 1])
    assert out[0, 20] == 2 and np.isnan(cm[0, 20])     # NaN < thr is false: accepted
    assert (out[0, 2:20] == 2).all() and (cm[0, 2:20] == 1.0).all()
    out_mv, cm_mv = oracle.agree(disp, L, R, thr=0.95, minvar=1.0, corrmap=True)
    assert out_mv[0, 20] == -32768 and cm_mv[0, 20] == -1.0   # variance filter rejects


# ------------------------------------------------------------ whole path

def test_match_recovers_ground_truth(oracle):
    """On a clean synthetic scene the matched integer disparities sit within
    1 px of the generator's ground truth (640x480 u8, 12 frames, FULL, 0.9)."""
    from libbicos_amd import scene
    L, R, gt = scene.generate(scene.SCENE_640x480_U8, seed=101, height=96)
    disp, _ = oracle.match(L, R, thr=0.9, full=True)
    v = disp != -32768
    assert v.mean() > 0.6
    assert (np.abs(disp[v] - gt[v]) <= 1.0).mean() > 0.95


def test_match_dispatch_and_errors(oracle):
    rng = np.random.default_rng(1)
    s = rng.integers(0, 256, size=(2, 4, 8)).astype(np.uint8)
    d, cm = oracle.match(s, s, thr=None)
    assert d.dtype == np.int16 and cm is None
    d, _ = oracle.match(s, s, thr=0.5, step=0.5)
    assert d.dtype == np.float32
    with pytest.raises(RuntimeError, match="at least two"):
        oracle.match(s[:1], s[:1])
    big = np.zeros((17, 2, 2), np.uint8)
    with pytest.raises(RuntimeError, match="258 bits"):
        oracle.match(big, big, full=True)
    assert oracle.required_bits(33, False) == 125 and oracle.descriptor_words(33, False) == 4
    assert oracle.required_bits(64, False) == 249 and oracle.descriptor_words(64, False) == 8
    assert oracle.required_bits(12, True) == 123


# --------------------------------------------------------- point clouds
# oracle/pointcloud.py restates cv::reprojectImageTo3D (src/cli.cpp:239) and
# save_pointcloud (include/fileutils.hpp:45-90); OpenCV is absent, so the
# restatement is pinned by closed forms.

def test_reproject_rectified_rig_closed_form():
    """Q of a rectified pair (focal f, principal point (cx, cy), baseline B,
    Tx = -B): the point of disparity d at (x, y) is ((x - cx), (y - cy), f) * B / d."""
    import pointcloud as PC
    f, cx, cy, B = 800.0, 319.5, 239.25, 0.12
    Q = [[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 0, f], [0, 0, 1 / B, 0]]
    rng = np.random.default_rng(11)
    d = rng.integers(1, 200, size=(6, 9)).astype(np.int16)
    p = PC.reproject(d, Q).astype(np.float64)
    y, x = np.mgrid[0:6, 0:9]
    s = B / d.astype(np.float64)
    exp = np.stack([(x - cx) * s, (y - cy) * s, f * s], axis=-1)
    assert np.allclose(p, exp, rtol=1e-6, atol=1e-6)
    # a flipped sign, a swapped x/y or a dropped translation fails this
    assert np.all(p[..., 2] > 0)


def test_reproject_identity_and_float_disparity():
    import pointcloud as PC
    d = np.array([[0.5, -3.25], [np.nan, 7.0]], np.float32)
    p = PC.reproject(d, np.eye(4))
    assert p[0, 1].tolist() == [1.0, 0.0, -3.25] and p[1, 1].tolist() == [1.0, 1.0, 7.0]
    assert np.isnan(p[1, 0]).all()                                   # 0 * NaN is NaN


def test_pointcloud_selection_hand_counted():
    """invalid disparities skipped silently; W = 0 gives non-finite points
    (counted); z < 0 counted unless allowed; row-major order."""
    import pointcloud as PC
    Q = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 1, 0]]   # W = d
    d = np.array([[2, -32768, 0], [-4, 5, 1]], np.int16)            # W = 0 at (0, 2)
    p = PC.reproject(d, Q)
    pts, n, nonfin, negz = PC.select(d, p)
    assert (n, nonfin, negz) == (4, 1, 0)                            # z = d / d = 1 wherever finite
    assert pts.tolist() == [[0.0, 0.0, 1.0], [0.0, -0.25, 1.0], [0.2, 0.2, 1.0], [2.0, 1.0, 1.0]] or \
        np.allclose(pts, [[0, 0, 1], [0, -0.25, 1], [0.2, 0.2, 1], [2, 1, 1]], rtol=1e-7)
    Qn = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]]  # z = -d, W = 1
    pts, n, nonfin, negz = PC.select(d, PC.reproject(d, Qn))
    assert (n, nonfin, negz) == (2, 0, 3)                            # z = -0.0 is not negative
    assert pts[:, 2].tolist() == [0.0, 4.0]
    pts, n, nonfin, negz = PC.select(d, PC.reproject(d, Qn), allow_negative_z=True)
    assert (n, nonfin, negz) == (5, 0, 0)
