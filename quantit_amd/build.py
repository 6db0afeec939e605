"""Build of libquantit_amd.so: every C++/HIP source compiled by hipcc for
gfx950 into quantit_amd/_build/, linked in-tree as quantit_amd/libquantit_amd.so.

Usage: python -m quantit_amd.build [--jobs N]
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
LIB = PKG / "libquantit_amd.so"
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled in every device object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(CSRC.glob("*.cpp")) + sorted((CSRC / "kernels").glob("*.hip"))


def compile_one(src: Path) -> Path:
    obj = BUILD / (src.stem + ("_k" if src.suffix == ".hip" else "") + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc(), *lang, *COMMON, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stderr}")
    os.replace(tmp, LIB)
    build_abi_client()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


ABI_CLIENT_SRC = PKG.parent / "tests" / "abi_client.c"
ABI_CLIENT = PKG.parent / "tests" / "abi_client"


def build_abi_client() -> Path:
    """The plain C client of include/quantit_amd.h (tests/abi_client.c),
    linked against the library with an rpath to it."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        raise RuntimeError("no C compiler for the ABI client")
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-I", str(PKG.parent / "include"), str(ABI_CLIENT_SRC), "-o",
           str(ABI_CLIENT), "-L", str(PKG), "-lquantit_amd", "-Wl,-rpath,$ORIGIN/../quantit_amd", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"ABI client build failed\n{res.stderr}")
    return ABI_CLIENT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
