"""bicos_hip_match_noalloc, the matching entry point that allocates nothing,
against bicos_hip_match through the C ABI: identical outputs bit for bit on
two calls in a row (the second on a smaller input, so stale contents of the
shared scratch from the first call are present), for every stage
combination; no device memory taken by its calls; inputs beyond its scratch
rejected before any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIGS = {
    "subpixel": dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0),
    "consistency": dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0,
                        variant=("cons", 1, False)),
    "moments": dict(nxcorr_threshold=0.96, subpixel_step=0.1, min_variance=2.0, precision="moments"),
    "agree_corrmap": dict(nxcorr_threshold=0.9, min_variance=2.0),
    "search_only": dict(nxcorr_threshold=None),
}


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_cfg(B, kw):
    kw = dict(kw)
    v = kw.pop("variant", None)
    if v is not None:
        kw["variant"] = B.Variant.Consistency(max_lr_diff=v[1], no_dupes=v[2])
    if kw.pop("precision", None) == "moments":
        kw["precision"] = B.Precision.Moments
    return B.Config(**kw)


@pytest.fixture(scope="module")
def pairs():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from libbicos_amd import scene, window
    L, R, _ = window.scene_device(scene.SCENE_1920x1200_U16, scene.BENCH_SEED)
    small = (L[:, 100:260, 300:940].contiguous(), R[:, 100:260, 300:940].contiguous())
    return (L, R), small


@pytest.mark.parametrize("name", list(CONFIGS))
def test_noalloc_equals_match_two_calls(pairs, name):
    """Full 1920x1200x33 pair, then a 160x640 crop: both calls equal the
    allocating entry point's outputs bit for bit."""
    import libbicos_amd as B
    cfg = make_cfg(B, CONFIGS[name])
    corr = name == "agree_corrmap"
    for L, R in pairs:
        exp, exp_cm = B.match_device(L, R, cfg, want_corrmap=corr)
        got, got_cm = B.match_device(L, R, cfg, want_corrmap=corr, noalloc=True)
        a, b = host(exp), host(got)
        assert a.dtype == b.dtype and np.array_equal(a, b, equal_nan=True), name
        if corr:
            assert np.array_equal(host(exp_cm), host(got_cm), equal_nan=True)


def test_noalloc_takes_no_device_memory(pairs):
    """After the first call on the device, calls into preallocated outputs
    leave the device's free memory unchanged (the allocating entry point's
    workspace is released first, so it cannot hide a new allocation)."""
    import torch
    import libbicos_amd as B
    (L, R), (l2, r2) = pairs
    cfg = make_cfg(B, CONFIGS["consistency"])
    out, _ = B.match_device(L, R, cfg, noalloc=True)
    out2, _ = B.match_device(l2, r2, cfg, noalloc=True)
    B.release_workspace()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        B.match_device(L, R, cfg, disparity=out, noalloc=True)
        B.match_device(l2, r2, cfg, disparity=out2, noalloc=True)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_rejects_beyond_scratch():
    """A 4096x3000 pair needs more scratch than the array holds: rejected
    with the reason, nothing launched; the allocating entry point serves it."""
    import torch
    import libbicos_amd as B
    cap = B.lib().bicos_hip_noalloc_scratch_bytes()
    assert cap == 96 << 20
    L = torch.zeros((2, 3000, 4096), dtype=torch.int16, device="cuda").view(torch.uint16)
    cfg = B.Config(nxcorr_threshold=0.9, subpixel_step=0.5)
    with pytest.raises(ValueError, match="scratch"):
        B.match_device(L, L, cfg, noalloc=True)
    d, _ = B.match_device(L, L, cfg)
    assert d.shape == (3000, 4096)
