"""Command-line front-end: the counterpart of the reference's bicos-cli
(src/cli.cpp:55-253, src/fileutils.cpp:59-170).

    python -m libbicos_amd.cli folder0 [folder1] [-t 0.75] [-v 1.0] [-s STEP]
        [-o bicosdisp.png] [-n N] [-q Q.yml] [--allow-negative-z]
        [-m LR_MAXDIFF] [--double] [--limited] [--corrmap] [--no-dupes]

Inputs: numbered PNG stacks, either folder0/N.png and folder1/N.png, or
folder0/N_left.png and folder0/N_right.png (sorted by N; 8/16-bit gray, colour
reduced to gray). As in the reference, the single-folder form reads 8-bit
gray (cv::IMREAD_GRAYSCALE without ANYDEPTH, src/fileutils.cpp:115) and the
two-folder form keeps 16-bit samples (ANYDEPTH, :73).

Outputs (src/fileutils.cpp:31-53, include/fileutils.hpp:45-90): the disparity
as a TURBO-colourised PNG and a raw TIFF (int16, or float32 with subpixel
refinement), the correlation map as <stem>-corrmap.{png,tiff} (VIRIDIS) with
--corrmap, and with -q the ASCII point cloud <stem>.xyz.

The stacks are uploaded to the GPU, matched there with BICOS::match
semantics, and the latencies of upload, match and download are printed like
the reference's CUDA build. Deviation: --allow-negative-z takes effect (the
reference registers the option but reads a different key, src/cli.cpp:233,
so there it never does).
"""
from __future__ import annotations

import argparse
import os
import re
import sys
import time

import numpy as np

from . import imageio as io

LICENSE_HEADER = ("libBICOS-compatible matching on AMD MI355X (libbicos_amd); "
                  "command-line front-end")


def parse_args(argv=None):
    p = argparse.ArgumentParser(prog="python -m libbicos_amd.cli",
                                description="BICOS stereo matching of two image stacks on the GPU")
    p.add_argument("folder0", help="folder of numbered input images (N.png, or N_left.png and "
                                   "N_right.png when folder1 is not given)")
    p.add_argument("folder1", nargs="?", default=None,
                   help="folder of the second view's images; with it both folders hold "
                        "0.png, 1.png, ...")
    p.add_argument("-t", "--threshold", type=float, default=0.75,
                   help="smallest normalized cross-correlation a match keeps (0 turns the test off)")
    # shown as default 1.0, but (like cxxopts' count(), which does not count a
    # default value) min_variance is only set when -v is given: None here
    p.add_argument("-v", "--variance", type=float, default=None,
                   help="smallest intensity variance of a matched pixel, with --threshold only "
                        "(default 1.0, applied only when the option is given)")
    p.add_argument("-s", "--step", type=float, default=None,
                   help="grid step of the subpixel refinement, with --threshold only")
    p.add_argument("-o", "--out", default="bicosdisp.png", help="disparity image to write")
    p.add_argument("-n", "--stacksize", type=int, default=None,
                   help="how many images per view to use (default: every one found)")
    p.add_argument("-q", "--qmatrix", default=None,
                   help="OpenCV FileStorage file holding the 4x4 matrix \"Q\"; writes a point cloud")
    p.add_argument("--allow-negative-z", action="store_true",
                   help="keep points with z < 0 in the point cloud")
    p.add_argument("-m", "--lr-maxdiff", type=int, default=None,
                   help="largest left/right disparity difference (the Consistency variant; "
                        "turns the duplicate filter off unless --no-dupes is given too)")
    p.add_argument("--double", action="store_true", help="FP64 correlation instead of FP32")
    p.add_argument("--limited", action="store_true",
                   help="limited descriptor transform (shorter descriptors, deeper stacks)")
    p.add_argument("--corrmap", action="store_true", help="also write the correlation map")
    p.add_argument("--no-dupes", action="store_true",
                   help="the duplicate filter (the variant used without --lr-maxdiff; with it, "
                        "both filters)")
    return p.parse_args(argv)


def config_from_args(args, B):
    """Option -> Config mapping of src/cli.cpp:137-166."""
    cfg = B.Config(nxcorr_threshold=args.threshold, mode=B.TransformMode.Full)
    if cfg.nxcorr_threshold is not None and cfg.nxcorr_threshold <= 0.0:
        cfg.nxcorr_threshold = None
    if args.corrmap and cfg.nxcorr_threshold is None:
        cfg.nxcorr_threshold = -1.0
        print(f"Computing with nxcorr-threshold of {cfg.nxcorr_threshold} because 'corrmap' is set",
              file=sys.stderr)
    if args.step is not None:
        cfg.subpixel_step = args.step
    if args.limited:
        cfg.mode = B.TransformMode.Limited
    if args.variance is not None and args.variance > 0.0:  # args.count("variance"), cli.cpp:154-156
        cfg.min_variance = args.variance
    if args.double:
        cfg.precision = B.Precision.Double
    if args.lr_maxdiff is not None:
        cfg.variant = B.Variant.Consistency(max_lr_diff=int(args.lr_maxdiff), no_dupes=bool(args.no_dupes))
    return cfg


_NUM = re.compile(r"^(\d+)")


def _index(name: str, errmsg: str) -> int:
    m = _NUM.match(name)
    if not m:
        raise ValueError(errmsg)
    return int(m.group(1))


def read_sequence(folder0: str, folder1=None):
    """Left and right stacks, each sorted by file number (src/fileutils.cpp:59-150)."""
    left, right = [], []
    if folder1 is not None:
        err = "Expecting numbered files with names NN.png; e.g 0.png, 1.png..."
        for folder, seq in ((folder0, left), (folder1, right)):
            for name in os.listdir(folder):
                seq.append((_index(name, err), io.to_gray(io.read_png(os.path.join(folder, name)))))
    else:
        err = "Expecting numbered files with names NN_{left,right}.png; e.g.: 5_left.png, 10_right.png..."
        for name in os.listdir(folder0):
            if "_" not in name:
                raise ValueError(err)
            img = io.to_gray(io.read_png(os.path.join(folder0, name)))
            if img.dtype == np.uint16:  # IMREAD_GRAYSCALE without ANYDEPTH: 8-bit
                img = (img >> 8).astype(np.uint8)
            (left if "_left" in name else right).append((_index(name, err), img))
    if len(left) != len(right):
        raise ValueError(f"Unequal number of images; left: {len(left)}, right: {len(right)}")
    left.sort(key=lambda e: e[0])
    right.sort(key=lambda e: e[0])
    return [m for _, m in left], [m for _, m in right]


def _with_ext(path: str, ext: str) -> str:
    return os.path.splitext(path)[0] + "." + ext


def save_image(image: np.ndarray, outfile: str, cmap: str = "turbo") -> None:
    """src/fileutils.cpp:31-53: colourised PNG of the valid range, raw TIFF."""
    valid = ~np.isnan(image) if image.dtype.kind == "f" else image != -32768
    rgb = io.colorize(io.normalize_minmax_u8(image, valid), valid, cmap)
    png = _with_ext(outfile, "png")
    io.write_png(png, rgb)
    print(f"Saved colorized disparity to\t\t{png}")
    tiff = _with_ext(outfile, "tiff")
    io.write_tiff(tiff, image)
    if image.dtype.kind == "f":
        print(f"Saved floating-point disparity to\t{tiff}")
    else:
        print(f"Saved raw disparity to\t\t\t{tiff}")


def save_pointcloud(points: np.ndarray, counts: dict, outfile: str) -> None:
    """include/fileutils.hpp:45-90: `x y z` per line, C++ ostream default
    formatting (6 significant digits)."""
    path = _with_ext(outfile, "xyz")
    with open(path, "w") as f:
        f.writelines(f"{x:g} {y:g} {z:g}\n" for x, y, z in points.tolist())
    print(f"Saved pointcloud in ascii-format to\t{path}")
    if counts["nonfinite"]:
        print(f"Skipped {counts['nonfinite']} points with non-finite fp values", file=sys.stderr)
    if counts["negative_z"]:
        print(f"Skipped {counts['negative_z']} points with negative Z values", file=sys.stderr)


def main(argv=None) -> int:
    args = parse_args(argv)
    import torch

    import libbicos_amd as B

    print(LICENSE_HEADER)
    if not sys.stdout.isatty():
        print("Danger: bicos-cli does not have a stable CLI interface", file=sys.stderr)
    if args.no_dupes and args.lr_maxdiff is None:
        print("'no-dupes' is the default when 'lr-maxdiff' is not set.", file=sys.stderr)
    q = None
    if args.qmatrix is not None:
        if not os.path.exists(args.qmatrix):
            raise ValueError(f"'{args.qmatrix}' does not exist")
        q = io.read_q_matrix(args.qmatrix)

    lstack, rstack = read_sequence(args.folder0, args.folder1)
    if args.stacksize is not None and args.stacksize < len(lstack):
        lstack, rstack = lstack[:args.stacksize], rstack[:args.stacksize]
    if not lstack:
        raise ValueError("no input images")
    print(f"Loaded {2 * len(lstack)} {8 * lstack[0].dtype.itemsize}-bit images in total")
    cfg = config_from_args(args, B)

    def dev(stack):
        a = np.ascontiguousarray(np.stack(stack))
        t = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).cuda()
        return t.view(torch.uint16) if a.dtype == np.uint16 else t

    torch.cuda.synchronize()
    tick = time.perf_counter()
    dl, dr = dev(lstack), dev(rstack)
    torch.cuda.synchronize()
    t_up = (time.perf_counter() - tick) * 1e3
    print(f"Latency:\t {t_up:.3f}ms (upload)\t", end="", flush=True)

    tick = time.perf_counter()
    disp_d, corr_d = B.match_device(dl, dr, cfg, want_corrmap=args.corrmap)
    torch.cuda.synchronize()
    t_match = (time.perf_counter() - tick) * 1e3
    print(f"{t_match:.3f}ms (match)\t", end="", flush=True)

    tick = time.perf_counter()
    disp = disp_d.cpu().numpy()
    corr = corr_d.cpu().numpy() if corr_d is not None else None
    t_down = (time.perf_counter() - tick) * 1e3
    print(f"{t_down:.3f}ms (download)")

    save_image(disp, args.out)
    if corr is not None:
        stem, ext = os.path.splitext(args.out)
        save_image(corr, stem + "-corrmap" + ext, "viridis")
    if q is not None:
        pts, counts = B.pointcloud(disp_d, q, allow_negative_z=args.allow_negative_z)
        save_pointcloud(pts.cpu().numpy(), counts, args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
