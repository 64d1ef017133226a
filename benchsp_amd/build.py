"""Build libbsp.so (HIP kernels for gfx950 + the C-ABI) in-tree.

Every source under benchsp_amd/csrc is compiled with hipcc into
benchsp_amd/_build/*.o and linked into benchsp_amd/libbsp.so. Objects are
rebuilt when their source or any header changed; `force=True` rebuilds all.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbsp.so")
ARCH = os.environ.get("BSP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    # No packed FP32 instructions in generated code: some of their forms have
    # produced wrong results on this hardware without any error.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
LINK = ["-shared", f"--offload-arch={ARCH}", "-lrocsparse", "-lpthread"]


def _sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers_mtime():
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    newest = max(os.path.getmtime(src), _headers_mtime())
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj + ".tmp"]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{r.stderr[-6000:]}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, *LINK, *objs, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


COMPAT_SRC = os.path.join(ROOT, "tests", "compat_driver.cpp")
COMPAT_BIN = os.path.join(ROOT, "tests", "compat_driver")


def build_compat_driver() -> str:
    """Compiles the C++ driver that uses include/benchsp_compat.hpp (the
    reference's call surface) and links it to libbsp.so."""
    tmp = COMPAT_BIN + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), COMPAT_SRC,
           "-L", HERE, "-lbsp", "-Wl,-rpath,$ORIGIN/../benchsp_amd",
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compat driver build failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, COMPAT_BIN)
    return COMPAT_BIN


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
    print(build_compat_driver())
