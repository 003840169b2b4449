"""Image files of the command-line front-end, without OpenCV.

The reference CLI reads its stacks with cv::imread and writes its results
with cv::imwrite (src/fileutils.cpp:31-170). This module covers the formats
that path uses:

- PNG read: 8/16-bit grayscale, grayscale+alpha, RGB, RGBA, non-interlaced
  (PNG specification, ISO/IEC 15948). Colour input is reduced to gray with
  the ITU-R BT.601 weights 0.299, 0.587, 0.114 (cv::imread's
  IMREAD_GRAYSCALE goes through libpng's own conversion, which may differ by
  one level); alpha is dropped (src/fileutils.cpp:83-87).
- PNG write: 8-bit RGB (colourised disparity) and 8/16-bit gray.
- TIFF write/read: uncompressed single-strip grayscale int16 / float32 /
  float64 (the raw disparity and correlation maps, src/fileutils.cpp:49).
- cv::normalize(NORM_MINMAX) to 8 bits over the valid pixels and the TURBO /
  VIRIDIS colour maps of save_image (src/fileutils.cpp:31-47). The colour maps
  are the published polynomial approximations (display only; the TIFF keeps
  the data).
- The Q matrix of a cv::FileStorage file (YAML or XML, `Q` as an
  opencv-matrix), as src/cli.cpp:234-238 reads it.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import struct
import zlib

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_io = None


def _iolib():
    global _io
    if _io is None:
        path = os.path.join(_HERE, "libbicos_io.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python libbicos_amd/build.py`")
        _io = C.CDLL(path)
        _io.bicos_png_unfilter.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int, C.c_void_p]
        _io.bicos_png_unfilter.restype = C.c_int
    return _io


PNG_SIG = b"\x89PNG\r\n\x1a\n"
_CHANNELS = {0: 1, 2: 3, 4: 2, 6: 4}


def read_png(path: str) -> np.ndarray:
    """A PNG file as an (H, W) array (gray) or (H, W, C) array, uint8 or
    uint16 (big-endian samples converted to native)."""
    data = open(path, "rb").read()
    if data[:8] != PNG_SIG:
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, ihdr = 8, [], None
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        ctype = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        (crc,) = struct.unpack(">I", data[pos + 8 + length:pos + 12 + length])
        if zlib.crc32(ctype + body) & 0xFFFFFFFF != crc:
            raise ValueError(f"{path}: CRC mismatch in chunk {ctype!r}")
        if ctype == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif ctype == b"IDAT":
            idat.append(body)
        elif ctype == b"IEND":
            break
        pos += 12 + length
    if ihdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, color, _comp, _filt, interlace = ihdr
    if color not in _CHANNELS or depth not in (8, 16):
        raise ValueError(f"{path}: unsupported PNG (colour type {color}, depth {depth})")
    if interlace:
        raise ValueError(f"{path}: interlaced PNG is not supported")
    ch = _CHANNELS[color]
    bpp = ch * depth // 8
    stride = w * bpp
    raw = zlib.decompress(b"".join(idat))
    if len(raw) != h * (stride + 1):
        raise ValueError(f"{path}: image data has the wrong size")
    src = np.frombuffer(raw, np.uint8)
    out = np.empty(h * stride, np.uint8)
    if _iolib().bicos_png_unfilter(src.ctypes.data, h, stride, bpp, out.ctypes.data) != 0:
        raise ValueError(f"{path}: unknown PNG filter type")
    img = out.view(">u2").astype(np.uint16) if depth == 16 else out
    img = img.reshape(h, w, ch)
    return img[:, :, 0] if ch == 1 else img


def to_gray(img: np.ndarray) -> np.ndarray:
    """Single-channel image: gray as is, alpha dropped, RGB by BT.601."""
    if img.ndim == 2:
        return img
    if img.shape[2] in (1, 2):
        return img[:, :, 0]
    rgb = img[:, :, :3].astype(np.float64)
    y = np.rint(rgb @ np.array([0.299, 0.587, 0.114]))
    return np.clip(y, 0, np.iinfo(img.dtype).max).astype(img.dtype)


def _chunk(ctype: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + ctype + body + struct.pack(">I", zlib.crc32(ctype + body) & 0xFFFFFFFF)


def write_png(path: str, img: np.ndarray) -> None:
    """(H, W) gray or (H, W, 3) RGB, uint8 or uint16; filter type 0 on every row."""
    img = np.ascontiguousarray(img)
    if img.dtype not in (np.uint8, np.uint16) or img.ndim not in (2, 3):
        raise ValueError("write_png takes uint8/uint16 gray or RGB images")
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    color = {1: 0, 3: 2}[ch]
    depth = 8 * img.dtype.itemsize
    rows = img.astype(">u2") if depth == 16 else img
    rows = rows.reshape(h, w * ch * img.dtype.itemsize // rows.dtype.itemsize).view(np.uint8).reshape(h, -1)
    filt = np.concatenate([np.zeros((h, 1), np.uint8), rows], axis=1)
    body = (PNG_SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 0))
            + _chunk(b"IDAT", zlib.compress(filt.tobytes(), 6)) + _chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(body)


_TIFF_FMT = {np.dtype(np.int16): (16, 2), np.dtype(np.uint16): (16, 1), np.dtype(np.uint8): (8, 1),
             np.dtype(np.float32): (32, 3), np.dtype(np.float64): (64, 3)}


def write_tiff(path: str, img: np.ndarray) -> None:
    """Uncompressed little-endian single-strip grayscale TIFF (baseline tags
    plus SampleFormat: 1 unsigned, 2 signed, 3 IEEE float)."""
    img = np.ascontiguousarray(img)
    if img.ndim != 2 or img.dtype not in _TIFF_FMT:
        raise ValueError("write_tiff takes 2-D int16/uint16/uint8/float32/float64 images")
    bits, fmt = _TIFF_FMT[img.dtype]
    h, w = img.shape
    pixels = img.astype(img.dtype.newbyteorder("<")).tobytes()
    tags = [(256, 4, w), (257, 4, h), (258, 3, bits), (259, 3, 1), (262, 3, 1), (273, 4, 0),
            (277, 3, 1), (278, 4, h), (279, 4, len(pixels)), (284, 3, 1), (339, 3, fmt)]
    ifd_off = 8
    data_off = ifd_off + 2 + 12 * len(tags) + 4
    ifd = struct.pack("<H", len(tags))
    for tag, typ, val in tags:
        if tag == 273:
            val = data_off
        ifd += struct.pack("<HHI", tag, typ, 1) + (struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val))
    ifd += struct.pack("<I", 0)
    with open(path, "wb") as f:
        f.write(b"II*\x00" + struct.pack("<I", ifd_off) + ifd + pixels)


def read_tiff(path: str) -> np.ndarray:
    """Reads back what write_tiff writes (uncompressed, one sample per pixel,
    strips in order; either byte order)."""
    data = open(path, "rb").read()
    bo = {b"II": "<", b"MM": ">"}.get(data[:2])
    if bo is None or struct.unpack(bo + "H", data[2:4])[0] != 42:
        raise ValueError(f"{path}: not a TIFF file")
    (off,) = struct.unpack(bo + "I", data[4:8])
    (n,) = struct.unpack(bo + "H", data[off:off + 2])
    tags = {}
    for i in range(n):
        tag, typ, cnt = struct.unpack(bo + "HHI", data[off + 2 + 12 * i:off + 10 + 12 * i])
        raw = data[off + 10 + 12 * i:off + 14 + 12 * i]
        size = {3: 2, 4: 4}.get(typ, 4)
        if cnt * size <= 4:
            vals = struct.unpack(bo + ("H" if typ == 3 else "I") * cnt, raw[:cnt * size])
        else:
            (p,) = struct.unpack(bo + "I", raw)
            vals = struct.unpack(bo + ("H" if typ == 3 else "I") * cnt, data[p:p + cnt * size])
        tags[tag] = vals
    if tags.get(259, (1,))[0] != 1:
        raise ValueError(f"{path}: compressed TIFF is not supported")
    w, h, bits = tags[256][0], tags[257][0], tags[258][0]
    fmt = tags.get(339, (1,))[0]
    dt = np.dtype({(8, 1): "u1", (16, 1): "u2", (16, 2): "i2", (32, 3): "f4", (64, 3): "f8",
                   (32, 2): "i4", (32, 1): "u4"}[(bits, fmt)]).newbyteorder(bo)
    buf = b"".join(data[o:o + c] for o, c in zip(tags[273], tags[279]))
    return np.frombuffer(buf, dt, count=w * h).reshape(h, w).astype(dt.newbyteorder("="))


def normalize_minmax_u8(img: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """cv::normalize(img, out, 0, 255, NORM_MINMAX, CV_8UC1, mask): the valid
    pixels' [min, max] mapped linearly onto [0, 255] (rounded, saturated);
    pixels outside the mask are 0 here (save_image zeroes them,
    src/fileutils.cpp:44)."""
    out = np.zeros(img.shape, np.uint8)
    if not valid.any():
        return out
    v = img[valid].astype(np.float64)
    lo, hi = v.min(), v.max()
    scale = 255.0 / (hi - lo) if hi > lo else 0.0
    out[valid] = np.clip(np.rint((v - lo) * scale), 0, 255).astype(np.uint8)
    return out


def _poly(x, coeffs):
    r = np.zeros_like(x)
    for c in coeffs[::-1]:
        r = r * x + c
    return r


# Polynomial fits of the colour maps (x in [0, 1] -> r, g, b in [0, 1]).
_TURBO = ((0.13572138, 4.61539260, -42.66032258, 132.13108234, -152.94239396, 59.28637943),
          (0.09140261, 2.19418839, 4.84296658, -14.18503333, 4.27729857, 2.82956604),
          (0.10667330, 12.64194608, -60.58204836, 110.36276771, -89.90310912, 27.34824973))
_VIRIDIS = ((0.2777273272234177, 0.1050930431085774, -0.3308618287255563, -4.634230498983486,
             6.228269936347081, 4.776384997670288, -5.435455855934631),
            (0.005407344544966578, 1.404613529898575, 0.214847559468213, -5.799100973351585,
             14.17993336680509, -13.74514537774601, 4.645852612178535),
            (0.3340998053353061, 1.384590162594685, 0.09509516302823659, -19.33244095627987,
             56.69055260068105, -65.35303263337234, 26.3124352495832))
COLORMAPS = {"turbo": _TURBO, "viridis": _VIRIDIS}


def colorize(u8: np.ndarray, valid: np.ndarray, cmap: str = "turbo") -> np.ndarray:
    """(H, W, 3) uint8 RGB of an 8-bit image through a colour map; invalid
    pixels black (src/fileutils.cpp:45-46)."""
    x = np.arange(256, dtype=np.float64) / 255.0
    lut = np.stack([_poly(x, c) for c in COLORMAPS[cmap]], axis=1)
    lut = np.clip(np.rint(lut * 255.0), 0, 255).astype(np.uint8)
    rgb = lut[u8]
    rgb[~valid] = 0
    return rgb


def read_q_matrix(path: str) -> np.ndarray:
    """The 4x4 matrix `Q` of a cv::FileStorage file (YAML `Q: !!opencv-matrix`
    or XML `<Q type_id="opencv-matrix">`), row-major float64."""
    text = open(path).read()
    m = re.search(r"(?:^|\n)\s*Q\s*:\s*!!opencv-matrix(.*?)(?:\n\S|\Z)", text, re.S)
    if m is None:
        m = re.search(r"<Q\b[^>]*>(.*?)</Q>", text, re.S)
        if m is None:
            raise ValueError(f"{path}: no matrix named Q")
        body = m.group(1)
        rows = int(re.search(r"<rows>\s*(\d+)", body).group(1))
        cols = int(re.search(r"<cols>\s*(\d+)", body).group(1))
        vals = re.search(r"<data>(.*?)</data>", body, re.S).group(1).split()
    else:
        body = m.group(1)
        rows = int(re.search(r"rows\s*:\s*(\d+)", body).group(1))
        cols = int(re.search(r"cols\s*:\s*(\d+)", body).group(1))
        vals = re.search(r"data\s*:\s*\[(.*?)\]", body, re.S).group(1).replace(",", " ").split()
    q = np.array([float(v) for v in vals], np.float64)
    if rows != 4 or cols != 4 or q.size != 16:
        raise ValueError(f"{path}: Q must be a 4x4 matrix")
    return q.reshape(4, 4)
