"""The C++ facade (include/ktplace_gpu.hpp) that a KTPlace build would use in
place of LinearSystems / solveBothAxes / LookAheadLegalizer: it compiles
against the C ABI header and links with the library (CPU), and on the GPU it
gives the same systems, solve and legalization as the Python binding of the
same library."""
import os
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "ktplace_amd"


def build_shim(out: Path) -> Path:
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    if not (LIBDIR / "libktplace_amd.so").exists():
        pytest.skip("library not built")
    exe = out / "shim_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "shim_check.cc"), "-o", str(exe), f"-L{LIBDIR}", "-lktplace_amd",
                    f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True, capture_output=True, timeout=300)
    return exe


def test_facade_compiles_and_links(tmp_path):
    exe = build_shim(tmp_path)
    assert exe.exists() and os.access(exe, os.X_OK)


def _bits(v: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', v))[0]:016x}"


@pytest.mark.gpu
def test_facade_matches_the_python_binding(tmp_path):
    from ktplace_amd import design as dz
    from ktplace_amd import simpl
    exe = build_shim(tmp_path)
    out = subprocess.run([str(exe), "3000", "3300", "60", "11"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = dict(line.split(" ", 1) for line in out.stdout.strip().split("\n"))
    d = dz.generate(3000, 3300, 60, seed=11)
    x, y = dz.seed_uniform(d, 20240607)
    with simpl.Context(0) as ctx:
        ctx.upload(d)
        ctx.build_b2b(x, y)
        sx, sy = ctx.get_system(0), ctx.get_system(1)
        r = ctx.solve_both(x, y, 1e-6, 100000)
        ctx.build_grid(0, 1.0)
        ctx.set_positions(x, y)
        lower = ctx.hpwl()
        ux, uy = ctx.lal(x, y)
        ctx.set_positions(ux, uy)
        upper, ovf = ctx.hpwl(), ctx.overflow()
    assert got["iterations"].split()[0] == str(r["iterations"])
    assert got["iterations"].split()[2] == str(r["iterations"])  # the observer saw every iterate
    assert got["iterations"].split()[4:] == [_bits(r["resid"][0]), _bits(r["resid"][1])]
    assert got["hpwl_lower"] == _bits(lower)
    assert got["hpwl_upper"] == _bits(upper) and got["overflow_upper"] == _bits(ovf)
    assert got["error"] == "raised"
    # the systems: the same hash of every word as the binding's copy
    def fnv(s):
        h = 1469598103934665603
        for arr in (s.row_ptr.astype(np.uint64), s.col.astype(np.uint64), s.val.view(np.uint64),
                    s.diag.view(np.uint64), s.rhs.view(np.uint64)):
            for v in arr.tolist():
                h = ((h ^ v) * 1099511628211) % (1 << 64)
        return f"{h:016x}"
    assert got["b2b_x"] == fnv(sx) and got["b2b_y"] == fnv(sy)
