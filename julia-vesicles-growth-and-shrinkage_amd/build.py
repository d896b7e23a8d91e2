"""Build recipe for libsav_hip.so (gfx950).  Compiles every csrc/*.hip with hipcc
in parallel into _build/ and links the shared library next to this file.

    python julia-vesicles-growth-and-shrinkage_amd/build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsav_hip.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("SAV_OFFLOAD_ARCH", "gfx950")
CFLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
          # packed FP32 instructions stay off: ordinary float code only
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stderr or "").splitlines() if "packed-fp32" not in l)
        raise RuntimeError(f"hipcc failed for {src.name}:\n{msg}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_client()
    if verbose:
        print(f"built {LIB}")
    return LIB


CLIENT_SRC = PKG / "abi_client.cpp"
CLIENT = PKG / "sav_abi_client"


def build_client() -> Path:
    """The plain C++ ABI client (host compiler only, linked against libsav_hip.so)."""
    cxx = shutil.which("g++") or "g++"
    tmp = CLIENT.with_suffix(".tmp")
    cmd = [cxx, "-O2", "-std=c++17", str(CLIENT_SRC), "-o", str(tmp), f"-L{PKG}", "-lsav_hip",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"abi client build failed:\n{r.stderr}")
    os.replace(tmp, CLIENT)
    return CLIENT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
