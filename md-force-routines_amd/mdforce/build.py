"""Builds libmdforce.so (all HIP sources, gfx950) in-tree.

    python -m mdforce.build        # from md-force-routines_amd/

Each ``csrc/*.hip`` is compiled to an object in ``_build/`` (in parallel)
and linked into ``mdforce/libmdforce.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG_DIR, "libmdforce.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    _build_cli(verbose)
    return LIB


CLI = os.path.join(ROOT, "mdf_bench")


def _build_cli(verbose: bool) -> None:
    """The standalone PP/PM timing program, a C caller of libmdforce."""
    src = os.path.join(CSRC, "cli", "mdf_bench.cpp")
    inc = os.path.join(os.path.dirname(ROOT), "include")
    tmp = CLI + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{inc}", src, "-o", tmp, f"-L{PKG_DIR}", "-lmdforce",
           "-Wl,-rpath,$ORIGIN/mdforce"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"mdf_bench build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {CLI}")


if __name__ == "__main__":
    build()
    sys.exit(0)
