"""Image I/O and option handling of the command-line front-end (CPU only).

PNG decoding is pinned by an encoder written here from the filter
definitions of the PNG specification (ISO/IEC 15948, section 9), a code path
independent of the decoder; TIFF by its tag layout; normalisation, gray
conversion and the Q-matrix reader by known answers; the option -> Config
mapping by src/cli.cpp:137-166.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from libbicos_amd import cli
from libbicos_amd import imageio as io


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _encode_png(img, filters):
    """PNG bytes with row r filtered by filters[r % len(filters)] (spec 9.2)."""
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    depth = 8 * img.dtype.itemsize
    raw = (img.astype(">u2") if depth == 16 else img).reshape(h, -1).view(np.uint8).reshape(h, -1)
    bpp = ch * depth // 8
    out = bytearray()
    for r in range(h):
        ft = filters[r % len(filters)]
        x = raw[r].astype(int)
        up = raw[r - 1].astype(int) if r else np.zeros_like(x)
        row = []
        for i in range(len(x)):
            a = x[i - bpp] if i >= bpp else 0
            b = up[i]
            c = up[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            row.append((x[i] - pred) & 0xFF)
        out += bytes([ft]) + bytes(row)
    color = {1: 0, 2: 4, 3: 2, 4: 6}[ch]

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)
    return (io.PNG_SIG + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("ch", [1, 2, 3, 4])
def test_png_decode_all_filters(tmp_path, dtype, ch):
    rng = np.random.default_rng(ch + 10 * (dtype == np.uint16))
    shape = (9, 13) if ch == 1 else (9, 13, ch)
    img = rng.integers(0, np.iinfo(dtype).max + 1, size=shape).astype(dtype)
    p = tmp_path / "f.png"
    p.write_bytes(_encode_png(img, [0, 1, 2, 3, 4, 4, 3, 2, 1]))
    assert np.array_equal(io.read_png(str(p)), img)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_png_write_read_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    for shape in ((31, 17), (8, 5, 3)):
        img = rng.integers(0, np.iinfo(dtype).max + 1, size=shape).astype(dtype)
        io.write_png(str(tmp_path / "r.png"), img)
        assert np.array_equal(io.read_png(str(tmp_path / "r.png")), img)


def test_png_rejects_bad_crc(tmp_path):
    data = bytearray(_encode_png(np.zeros((2, 2), np.uint8), [0]))
    data[-20] ^= 0xFF   # inside the IDAT payload
    (tmp_path / "bad.png").write_bytes(bytes(data))
    with pytest.raises(ValueError):
        io.read_png(str(tmp_path / "bad.png"))


def test_to_gray_bt601_known_answers():
    rgb = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255]]], np.uint8)
    assert io.to_gray(rgb).tolist() == [[76, 150, 29, 255]]
    ga = np.array([[[7, 200]]], np.uint16)
    assert io.to_gray(ga).tolist() == [[7]]


@pytest.mark.parametrize("dtype,fmt", [(np.int16, 2), (np.float32, 3), (np.float64, 3), (np.uint16, 1)])
def test_tiff_round_trip_and_tags(tmp_path, dtype, fmt):
    rng = np.random.default_rng(2)
    img = (rng.normal(0, 300, size=(7, 11))).astype(dtype)
    if dtype in (np.float32, np.float64):
        img[0, 0] = np.nan
    p = str(tmp_path / "t.tiff")
    io.write_tiff(p, img)
    got = io.read_tiff(p)
    assert got.dtype == img.dtype and np.array_equal(got, img, equal_nan=img.dtype.kind == "f")
    data = open(p, "rb").read()
    assert data[:4] == b"II*\x00"
    n = struct.unpack("<H", data[8:10])[0]
    tags = {struct.unpack("<H", data[10 + 12 * i:12 + 12 * i])[0]: data[18 + 12 * i:22 + 12 * i] for i in range(n)}
    assert struct.unpack("<H", tags[339][:2])[0] == fmt
    assert struct.unpack("<H", tags[258][:2])[0] == 8 * img.dtype.itemsize


def test_normalize_minmax_known_answers():
    img = np.array([[-32768, 10, 20], [30, 40, -32768]], np.int16)
    valid = img != -32768
    assert io.normalize_minmax_u8(img, valid).tolist() == [[0, 0, 85], [170, 255, 0]]
    f = np.array([[np.nan, 1.0, 1.0]], np.float32)
    assert io.normalize_minmax_u8(f, ~np.isnan(f)).tolist() == [[0, 0, 0]]
    rgb = io.colorize(np.array([[0, 255]], np.uint8), np.array([[False, True]]))
    assert rgb[0, 0].tolist() == [0, 0, 0] and rgb.shape == (1, 2, 3)


Q = [[1, 0, 0, -320.5], [0, 1, 0, -240.25], [0, 0, 0, 500.0], [0, 0, 0.1, 0]]


def test_q_matrix_yaml_and_xml(tmp_path):
    y = tmp_path / "q.yml"
    y.write_text("%YAML:1.0\n---\nQ: !!opencv-matrix\n   rows: 4\n   cols: 4\n   dt: d\n"
                 "   data: [ 1., 0., 0., -3.2050000000000000e+02, 0., 1., 0., -240.25, 0., 0., 0.,\n"
                 "       500., 0., 0., 1.0000000000000001e-01, 0. ]\n")
    assert np.allclose(io.read_q_matrix(str(y)), Q, rtol=0, atol=0)
    x = tmp_path / "q.xml"
    x.write_text('<?xml version="1.0"?>\n<opencv_storage>\n<Q type_id="opencv-matrix">\n'
                 "  <rows>4</rows>\n  <cols>4</cols>\n  <dt>d</dt>\n  <data>\n"
                 "    1. 0. 0. -320.5 0. 1. 0. -240.25 0. 0. 0. 500. 0. 0. 0.1 0.</data></Q>\n"
                 "</opencv_storage>\n")
    assert np.array_equal(io.read_q_matrix(str(x)), np.array(Q))


def test_cli_option_mapping():
    import libbicos_amd as B
    # min_variance only when -v is given (src/cli.cpp:154-156: cxxopts count()
    # does not count the option's default value)
    c = cli.config_from_args(cli.parse_args(["f"]), B)
    assert (c.nxcorr_threshold, c.min_variance, c.subpixel_step, c.mode) == (0.75, None, None, B.TransformMode.Full)
    c = cli.config_from_args(cli.parse_args(["f", "-v", "1.0"]), B)
    assert c.min_variance == 1.0
    c = cli.config_from_args(cli.parse_args(["f", "--variance", "2.5"]), B)
    assert c.min_variance == 2.5
    assert isinstance(c.variant, B.Variant.NoDuplicates) and c.precision == B.Precision.Single
    c = cli.config_from_args(cli.parse_args(["f", "-t", "0", "-v", "0"]), B)
    assert c.nxcorr_threshold is None and c.min_variance is None
    c = cli.config_from_args(cli.parse_args(["f", "-t", "0", "--corrmap"]), B)
    assert c.nxcorr_threshold == -1.0
    c = cli.config_from_args(cli.parse_args(["f", "g", "-s", "0.1", "--limited", "--double", "-m", "2", "--no-dupes"]), B)
    assert c.subpixel_step == 0.1 and c.mode == B.TransformMode.Limited and c.precision == B.Precision.Double
    assert isinstance(c.variant, B.Variant.Consistency) and (c.variant.max_lr_diff, c.variant.no_dupes) == (2, True)
    c = cli.config_from_args(cli.parse_args(["f", "-m", "1"]), B)
    assert c.variant.no_dupes is False


def test_read_sequence_orders_numerically(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir(), b.mkdir()
    for i in (10, 2, 1):
        io.write_png(str(a / f"{i}.png"), np.full((3, 4), i, np.uint16))
        io.write_png(str(b / f"{i}.png"), np.full((3, 4), 100 + i, np.uint16))
    left, right = cli.read_sequence(str(a), str(b))
    assert [int(m[0, 0]) for m in left] == [1, 2, 10] and left[0].dtype == np.uint16
    assert [int(m[0, 0]) for m in right] == [101, 102, 110]
    s = tmp_path / "s"
    s.mkdir()
    for i in (3, 1):
        io.write_png(str(s / f"{i}_left.png"), np.full((2, 2), 256 * i + 7, np.uint16))
        io.write_png(str(s / f"{i}_right.png"), np.full((2, 2), 256 * (i + 1), np.uint16))
    left, right = cli.read_sequence(str(s))
    assert [int(m[0, 0]) for m in left] == [1, 3] and left[0].dtype == np.uint8   # high byte
    assert [int(m[0, 0]) for m in right] == [2, 4]
    (s / "x.png").write_bytes(b"")
    with pytest.raises(ValueError):
        cli.read_sequence(str(s))
