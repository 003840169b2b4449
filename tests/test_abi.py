"""C ABI and host-logic checks that need no GPU: the library loads, exports
every function include/bicos_hip.h declares, reports the reference's type
rules, and rejects invalid calls before touching the device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "bicos_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bicos_hip_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import libbicos_amd as B
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(B.lib(), name), name
    assert set(B.EXPORTED_SYMBOLS) <= set(names)


def test_config_defaults_match_reference():
    """include/common.hpp:73-82"""
    import libbicos_amd as B
    c = B._Config()
    B.lib().bicos_hip_config_default(C.byref(c))
    assert c.has_nxcorr_threshold == 1 and abs(c.nxcorr_threshold - 0.5) < 1e-7
    assert c.has_subpixel_step == 0 and c.has_min_variance == 0
    assert c.mode == 0 and c.precision == 0 and c.variant == 0
    assert c.max_lr_diff == 1 and c.no_dupes == 0
    py = B.Config()
    assert py.nxcorr_threshold == 0.5 and py.subpixel_step is None and py.min_variance is None
    assert py.mode == B.TransformMode.Limited and isinstance(py.variant, B.Variant.NoDuplicates)
    v = B.Variant.Consistency()
    assert v.max_lr_diff == 1 and v.no_dupes is False


def test_output_type_rules():
    """int16 unless threshold and step are both set (cpu.cpp:85-88); corrmap
    F32/F64 only with a threshold (cuda.cu.in:183-191)."""
    import libbicos_amd as B
    assert B.disparity_dtype(B.Config()) == np.int16
    assert B.disparity_dtype(B.Config(subpixel_step=0.1)) == np.float32
    assert B.disparity_dtype(B.Config(nxcorr_threshold=None, subpixel_step=0.1)) == np.int16
    assert B.corrmap_dtype(B.Config()) == np.float32
    assert B.corrmap_dtype(B.Config(precision=B.Precision.Double)) == np.float64
    assert B.corrmap_dtype(B.Config(nxcorr_threshold=None)) is None
    assert B.invalid_disparity(np.int16) == -32768
    assert np.isnan(B.invalid_disparity(np.float32))
    with pytest.raises(B.Exception):
        B.invalid_disparity(np.uint8)


def test_descriptor_width_table():
    """src/impl/cpu.cpp:118-152"""
    import libbicos_amd as B
    table = {(2, 0): 32, (9, 0): 32, (10, 0): 64, (17, 0): 64, (18, 0): 128, (33, 0): 128,
             (34, 0): 256, (65, 0): 256, (66, 0): -1, (6, 1): 32, (7, 1): 64, (8, 1): 64,
             (9, 1): 128, (12, 1): 128, (13, 1): 256, (16, 1): 256, (17, 1): -1}
    for (n, mode), bits in table.items():
        assert B.descriptor_bits(n, mode) == bits, (n, mode)
    import oracle as Or  # the same table in the oracle
    for (n, mode), bits in table.items():
        w = Or.descriptor_words(n, bool(mode))
        assert (32 * w if w else -1) == bits


def _fake_stack(n, rows=4, cols=8, t=0):
    import libbicos_amd as B
    buf = (C.c_uint8 * (rows * cols * 2))()
    imgs = (B._Image * max(n, 1))(*[B._Image(C.addressof(buf), rows, cols, t, 0, cols * 2)
                                    for _ in range(max(n, 1))])
    imgs._buf = buf
    return imgs


def _call_match(n, cfg, t=0, disp_type=3):
    import libbicos_amd as B
    s = _fake_stack(n, t=t)
    buf = (C.c_uint8 * 256)()
    d = B._Image(C.addressof(buf), 4, 8, disp_type, 0, 32)
    rc = B.lib().bicos_hip_match(s, s, n, C.byref(cfg._c()), C.byref(d), None, None)
    return rc, B.lib().bicos_hip_last_error().decode()


def test_host_validation_errors():
    """Errors the reference raises before any device work (cpu.cpp:103-151,
    cuda.cu.in:241-289), reported as codes without touching the GPU."""
    import libbicos_amd as B
    rc, msg = _call_match(1, B.Config())
    assert rc == B.ERR_EXCEPTION and "at least two" in msg
    rc, msg = _call_match(4, B.Config(), t=5)
    assert rc == B.ERR_EXCEPTION and "bad input depths" in msg
    rc, msg = _call_match(17, B.Config(mode=B.TransformMode.Full))
    assert rc == B.ERR_INVALID_ARGUMENT and "258 bits" in msg
    rc, msg = _call_match(4, B.Config(subpixel_step=0.0), disp_type=5)
    assert rc == B.ERR_INVALID_ARGUMENT and "subpixel_step" in msg
    rc, msg = _call_match(4, B.Config(subpixel_step=0.1), disp_type=3)
    assert rc == B.ERR_EXCEPTION and "wrong element type" in msg
    rc, msg = _call_match(66, B.Config())
    assert rc == B.ERR_EXCEPTION and "Bad number of images" in msg


def test_cpp_headers_compile():
    """include/bicos/match.hpp is plain C++17 (no HIP, no OpenCV needed)."""
    src = '#include "bicos/match.hpp"\nint main(){ BICOS::Config c; (void)c; return 0; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include", "-x", "c++", "-"],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", f"-I{ROOT}/include", "-x", "c", "-"],
                       input='#include "bicos_hip.h"\nint main(void){return 0;}\n', text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


def test_scene_generator_is_deterministic():
    from libbicos_amd import scene
    a = scene.generate(scene.SCENE_640x480_U8, seed=5, height=24, width=100, threads=1)
    b = scene.generate(scene.SCENE_640x480_U8, seed=5, height=24, width=100, threads=4)
    c = scene.generate(scene.SCENE_640x480_U8, seed=6, height=24, width=100)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    L, R, gt = a
    assert L.shape == (12, 24, 100) and L.dtype == np.uint8
    assert 8.0 <= gt.min() and gt.max() <= 72.0
    u = scene.generate(scene.SCENE_1920x1200_U16, seed=5, height=8, width=300)[0]
    assert u.dtype == np.uint16 and u.max() > 40000


def test_window_host_side():
    """Learned-window entry points that run on the host: defaults, the
    deterministic initial parameters and argument checks before device work."""
    import libbicos_amd as B
    from libbicos_amd import window as Wn
    import cnn_ref
    assert Wn.PARAM_COUNT == cnn_ref.PARAMS == 28928
    o = B._WindowOptions()
    B.lib().bicos_hip_window_options_default(C.byref(o))
    py = Wn.WindowOptions()
    assert (o.in_scale, o.p_conf, o.hw_lo, o.hw_hi, o.bins, o.fallback) == \
        (py.in_scale, py.p_conf, py.hw_lo, py.hw_hi, py.bins, int(py.fallback))
    a, b, c = Wn.init_params(3), Wn.init_params(3), Wn.init_params(4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for l in range(4):  # He-normal weights, zero biases
        w = a[cnn_ref.w_off(l):cnn_ref.b_off(l)]
        fan_in = (4 if l == 0 else 32) * 9
        assert abs(w.std() * np.sqrt(fan_in / 2) - 1) < 0.1
        assert not a[cnn_ref.b_off(l):cnn_ref.b_off(l) + 32].any()
    s = _fake_stack(3)
    cfg = B.Config()._c()
    d = B._Image(None, 4, 8, 3, 0, 16)
    rc = B.lib().bicos_hip_match_windowed(s, s, 3, C.byref(cfg), C.c_void_p(1), C.byref(o),
                                          C.byref(d), None, None, None)
    assert rc == B.ERR_INVALID_ARGUMENT and "4 frames" in B.lib().bicos_hip_last_error().decode()
    o.bins = 3
    s = _fake_stack(4)
    rc = B.lib().bicos_hip_match_windowed(s, s, 4, C.byref(cfg), C.c_void_p(1), C.byref(o),
                                          C.byref(d), None, None, None)
    assert rc == B.ERR_INVALID_ARGUMENT and "bins" in B.lib().bicos_hip_last_error().decode()


def test_window_weights_roundtrip(tmp_path):
    from libbicos_amd import window as Wn
    p = Wn.init_params(9)
    f = str(tmp_path / "w.npz")
    Wn.save(f, p, {"scenes": 12, "note": "x"})
    q, meta = Wn.load(f)
    assert np.array_equal(p, q) and meta == {"scenes": "12", "note": "'x'"}


def test_search_chunk_cols():
    """The full-row search's LDS chunk sizes (host logic, no device needed):
    a 1920-column row of 128-bit descriptors is two fp4 chunks."""
    import libbicos_amd as B
    assert 2 * B.search_chunk_cols(B.SEARCH_MFMA, 4) == 1920
    assert B.search_chunk_cols(B.SEARCH_VALU, 4) == 0
    for w in (1, 2, 4, 8):
        assert B.search_chunk_cols(B.SEARCH_MFMA_I8, w) > 0 and B.search_chunk_cols(B.SEARCH_MFMA, w) > 0
    with pytest.raises(ValueError):
        B.search_chunk_cols(B.SEARCH_MFMA, 3)
