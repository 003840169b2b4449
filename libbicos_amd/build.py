"""Build recipe of the native parts of libbicos_amd.

    python libbicos_amd/build.py        # (a plain script: it must not import the
                                        #  package, whose import loads the library)

Produces, in-tree next to this file:
  libbicos_hip.so    gfx950 kernels + C ABI (include/bicos_hip.h) + C++ API
                     (include/bicos/match.hpp)
  libbicos_scene.so  synthetic scene generator (host C++)
  libbicos_io.so     PNG unfiltering for the command-line front-end (host C++)
  bicos_cli          small C++ program exercising the C++ API (tests)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: the reference CPU arithmetic forms no FMAs beyond the
# explicit fmaf calls; packed FP32 ops are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", f"-I{INCLUDE}",
]

HIP_SOURCES = ["bicos_hip.hip", "window.hip", "scene_gpu.hip", "pointcloud.hip", "match.cpp"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_hip() -> str:
    out = os.path.join(HERE, "libbicos_hip.so")
    objs, procs = [], []
    for src in HIP_SOURCES:  # translation units compile in parallel
        obj = os.path.join(CSRC, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = None
    for cmd, pr in procs:
        text, _ = pr.communicate()
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if text.strip():
            print(text, flush=True)
        if pr.returncode != 0 and failed is None:
            failed = cmd
    if failed:
        raise RuntimeError(f"build step failed: {' '.join(failed)}")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def build_scene() -> str:
    out = os.path.join(HERE, "libbicos_scene.so")
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread",
          "-Wall", "-o", out + ".tmp", os.path.join(CSRC, "scene.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_io() -> str:
    out = os.path.join(HERE, "libbicos_io.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", out + ".tmp",
          os.path.join(CSRC, "imageio.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_cli() -> str:
    out = os.path.join(HERE, "bicos_cli")
    _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", out + ".tmp",
          os.path.join(CSRC, "cli.cpp"), f"-L{HERE}", "-lbicos_hip", f"-Wl,-rpath,{HERE}",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build() -> None:
    build_scene()
    build_io()
    build_hip()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)
