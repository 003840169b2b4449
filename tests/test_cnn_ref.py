"""Pins the numpy reference of the learned-window CNN (oracle/cnn_ref.py):
its analytic gradient against central finite differences, its correlation
against a direct restatement, and its loss against a closed form."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref():
    import cnn_ref
    return cnn_ref


def tiny_scene(seed, H=16, W=96):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 65536, size=(4, H, W + 40)).astype(np.float64)
    d = 8                                  # 2 cells of disparity
    L = base[:, :, 20:20 + W]
    R = base[:, :, 20 + d:20 + d + W]      # left x sees right x - d... shifted content
    label = np.full((H // 4, W // 4), float(d))
    label[0, :3] = np.nan
    return L.astype(np.uint16), R.astype(np.uint16), label


def test_gradient_matches_finite_differences(ref):
    L, R, label = tiny_scene(1)
    p = ref.init_params(3).astype(np.float64)
    scale = 16 / 65535
    loss, g, _ = ref.loss_and_grad(L, R, label, p, scale, bf=False)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.integers(ref.w_off(l), ref.b_off(l) + ref.C, 6) for l in range(4)])
    eps = 1e-4
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (ref.loss_and_grad(L, R, label, pp, scale, bf=False)[0]
              - ref.loss_and_grad(L, R, label, pm, scale, bf=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 + 1e-3 * abs(fd), (i, fd, g[i])


def test_correlation_direct(ref):
    rng = np.random.default_rng(2)
    fl = rng.normal(size=(32, 3, 30))
    fr = rng.normal(size=(32, 3, 30))
    Cv = ref.corr(fl, fr)
    for (k, Y, X) in [(0, 0, 0), (8, 1, 5), (20, 2, 29), (79, 0, 29)]:
        d = ref.D0 + k
        if 0 <= X - d < 30:
            assert np.isclose(Cv[k, Y, X], fl[:, Y, X] @ fr[:, Y, X - d] / np.sqrt(32))
        else:
            assert Cv[k, Y, X] < -1e29


def test_softmax_ce_closed_form(ref):
    """uniform logits over the valid bins: loss = log(#valid); a label half
    way between two bins splits the target evenly"""
    Kn = ref.K
    Cv = np.zeros((Kn, 1, 40))
    Cv[:10] = -1e30
    lab = np.full((1, 40), np.nan)
    lab[0, 30] = 4.0 * (ref.D0 + 20.5)       # bins 20, 21: right cells 18, 17 exist
    lab[0, 5] = 4.0 * (ref.D0 + 20.0)        # bin 20 would need right cell -7: unlabelled
    loss, cnt, dC = ref.softmax_ce(Cv, lab)
    assert cnt == 1 and np.isclose(loss, np.log(Kn - 10))
    assert np.isclose(dC[20, 0, 30], 1 / (Kn - 10) - 0.5) and np.isclose(dC[21, 0, 30], 1 / (Kn - 10) - 0.5)
    assert np.isclose(dC[:, 0, 30].sum(), 0.0) and not dC[:, 0, 5].any()


def test_bf16_rounding(ref):
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0], np.float32)
    r = ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                 # exact tie -> even
    assert r[2] == np.float32(1.0078125)
    assert np.isclose(r[3], -3.0e-3, rtol=4e-3)
    assert r[4] == 65536.0
