"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

- The CPU oracle, the scene generator and the PNG unfiltering, built with
  g++ -fsanitize=address,undefined together with tests/sanitize/host_checks.cpp.
- The host side of libbicos_hip (argument validation, type and width rules,
  option defaults, the network's host initialisation), its sources built with
  hipcc with the sanitizers on the host code only (-Xarch_host) and linked with
  tests/sanitize/hip_host_checks.cpp; no call of the driver reaches the device.

Any sanitizer report aborts the program (-fno-sanitize-recover=all), so a
zero exit status and the driver's final line are the pass condition.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libbicos_amd", "csrc")
SAN_DIR = os.path.join(ROOT, "tests", "sanitize")
HIPCC = "/opt/rocm/bin/hipcc"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_oracle_scene_png_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_checks")
    _run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
          "-pthread", "-o", exe, os.path.join(SAN_DIR, "host_checks.cpp"),
          os.path.join(ROOT, "oracle", "bicos_oracle.cpp"), os.path.join(CSRC, "scene.cpp"),
          os.path.join(CSRC, "imageio.cpp")], timeout=600)
    r = _run([exe], env=ENV, timeout=300)
    assert r.stdout.strip().endswith("host checks ok")


@pytest.mark.timeout(900)
@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_hip_library_host_side_under_sanitizers(tmp_path):
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all"]
    inc = f"-I{os.path.join(ROOT, 'include')}"
    srcs = ["bicos_hip.hip", "window.hip", "pointcloud.hip", "scene_gpu.hip"]

    def compile_one(src):
        obj = str(tmp_path / (src + ".o"))
        # device code at -O1: only the host side runs here
        _run([HIPCC, "--offload-arch=gfx950", "-x", "hip", "-O1", "-std=c++17", "-fPIC", *san, inc,
              "-c", os.path.join(CSRC, src), "-o", obj], timeout=900)
        return obj

    with ThreadPoolExecutor(max_workers=4) as pool:
        objs = list(pool.map(compile_one, srcs))
    drv = str(tmp_path / "drv.o")
    _run([HIPCC, "-O1", "-g", "-std=c++17", *san, inc, "-c", os.path.join(SAN_DIR, "hip_host_checks.cpp"),
          "-o", drv], timeout=300)
    exe = str(tmp_path / "hip_host_checks")
    _run([HIPCC, "--offload-arch=gfx950", "-fsanitize=address,undefined", "-o", exe, drv, *objs], timeout=300)
    r = _run([exe], env=ENV, timeout=300)
    assert r.stdout.strip().endswith("hip host checks ok")
