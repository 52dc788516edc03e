"""The product's host geometry of look-ahead legalization (csrc/lal_geom.cc:
free space, the grid's free area per bin, free-area profiles) against the
oracle's restatement of simpl_grid.h:67-231 and simpl_placer.cc:190-224, bit
for bit, on the CPU.  lal_geom.cc is compiled with g++ into a small harness
(tests/lal_geom_check.cc); the GPU legalizer's parity tests cover the rest."""
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "ktplace_amd" / "csrc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("geom") / "lal_geom_check"
    subprocess.run([gxx, "-O2", "-std=c++20", "-ffp-contract=off", f"-I{CSRC}",
                    str(ROOT / "tests" / "lal_geom_check.cc"), str(CSRC / "lal_geom.cc"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    return exe


def _bits(a) -> list:
    return [struct.unpack("<Q", struct.pack("<d", float(v)))[0] for v in a]


def _run(exe, d, bins, boxes) -> tuple:
    raw, flat = d.raw, d.flat
    lines = [str(len(raw.row_y))]
    for r in range(len(raw.row_y)):
        lines.append(" ".join(float(v).hex() for v in (raw.row_y[r], raw.row_height[r], raw.row_x0[r],
                                                          raw.row_x1[r], raw.row_pitch[r])))
    lines.append(" ".join(float(v).hex() for v in raw.die))
    obs = np.asarray(flat.obstacles, np.float64).reshape(-1, 4)
    lines.append(str(len(obs)))
    lines += [" ".join(float(v).hex() for v in o) for o in obs]
    lines.append(str(bins))
    lines.append(str(len(boxes)))
    lines += [" ".join(float(v).hex() for v in b[:4]) + f" {int(b[4])}" for b in boxes]
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True, timeout=120).stdout.split("\n")
    avail = [int(t, 16) for t in out[0].split()[1:]]
    profiles = []
    for k in range(len(boxes)):
        e = [int(t, 16) for t in out[1 + 2 * k].split()[1:]]
        p = [int(t, 16) for t in out[2 + 2 * k].split()[1:]]
        profiles.append((e, p))
    return avail, profiles


@pytest.mark.parametrize("seed,bins,macros", [(9, 0, 0), (4, 37, 12), (5, 16, 30)])
def test_free_space_grid_and_profiles_match_the_oracle(harness, seed, bins, macros):
    d = dz.generate(3000, 3300, 60, seed=seed, num_macros=macros)  # macros are cut out of the rows
    assert len(np.asarray(d.flat.obstacles).reshape(-1, 4)) == macros
    o = ol.Oracle.of(d)
    per = o.build_grid(bins, 1.0)
    rng = np.random.default_rng(seed)
    W, H = d.die[2], d.die[3]
    boxes = []
    for k in range(200):
        x0, y0 = rng.uniform(0, W), rng.uniform(0, H)
        x1 = x0 + rng.uniform(0, W - x0) * (0.1 if k % 3 == 0 else 1.0)
        y1 = y0 + rng.uniform(0, H - y0) * (0.1 if k % 3 == 1 else 1.0)
        boxes.append((x0, y0, x1, y1, k % 2))
    boxes.append((0.0, 0.0, W, H, 1))
    boxes.append((0.0, 0.0, W, H, 0))
    avail, profiles = _run(harness, d, per, boxes)
    ref_avail, _ = o.grid_avail()
    assert avail == _bits(ref_avail)
    for b, (e, p) in zip(boxes, profiles):
        re, rp = o.profile(b[:4], bool(b[4]))
        assert e == _bits(re) and p == _bits(rp), b
