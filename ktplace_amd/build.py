"""In-tree build of libktplace_amd.so (HIP kernels for gfx950 + host C++).

Every source under ktplace_amd/csrc is compiled with hipcc for gfx950 only and
linked into ktplace_amd/libktplace_amd.so.  Device code is built with
-ffp-contract=off (the B2B assembly must round exactly like the reference's
-ffp-contract=off build; fused multiply-adds are written out explicitly where
a kernel wants them) and without packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
BUILD = PKG / "_build"
LIB = PKG / "libktplace_amd.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    f"-I{INCLUDE}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every HIP/C++ source and link the shared library (always rebuilds)."""
    if BUILD.exists():
        shutil.rmtree(BUILD)
    BUILD.mkdir(parents=True)
    sources = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cc")])
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(jobs, 1)) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz",
           "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    oracle = PKG.parent / "oracle"
    subprocess.run(["make", "-B", "-C", str(oracle)], check=True, capture_output=True)
    return oracle / "liboracle.so"


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
