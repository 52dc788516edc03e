"""The product's host geometry of look-ahead legalization (csrc/lal_geom.cc:
free space, the grid's free area per bin, free-area profiles, stripes in
fill order, overfilled clusters) against the oracle's restatement of simpl_grid.h:67-231,
simpl_placer.cc:190-224 and simpl_lal.cc:238-282, bit for bit, on the CPU.  lal_geom.cc is compiled with g++ into a small harness
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


def _run(exe, d, bins, queries) -> tuple:
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
    lines.append(str(len(queries)))
    for q in queries:  # ("P", box, vertical), ("S", box, vertical, cell, room, g, frac, cut, below)
        if q[0] == "C":  # ("C", density limit, occupancy per bin)
            lines.append("2 " + float(q[1]).hex() + " " + " ".join(float(v).hex() for v in q[2]))
            continue
        nums = list(q[1]) + [int(q[2])] + [v for v in q[3:]]
        kind = {"P": "0 ", "S": "1 ", "p": "3 ", "s": "4 "}[q[0]]
        lines.append(kind +
                     " ".join(v.hex() if isinstance(v, float) else str(v) for v in map(_num, nums)))
    out = iter(subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                              check=True, timeout=120).stdout.split("\n"))
    avail = [int(t, 16) for t in next(out).split()[1:]]
    results = []
    for q in queries:
        if q[0] in "Ss":
            results.append([int(t, 16) for t in next(out).split()[1:]])
        elif q[0] == "C":
            results.append([int(t) for t in next(out).split()[1:]])
        else:
            e = [int(t, 16) for t in next(out).split()[1:]]
            p = [int(t, 16) for t in next(out).split()[1:]]
            results.append((e, p))
    return avail, results


def _num(v):
    return int(v) if isinstance(v, (bool, np.bool_)) else float(v)


@pytest.mark.parametrize("seed,bins,macros", [(9, 0, 0), (4, 37, 12), (5, 16, 30)])
def test_free_space_grid_and_profiles_match_the_oracle(harness, seed, bins, macros):
    d = dz.generate(3000, 3300, 60, seed=seed, num_macros=macros)  # macros are cut out of the rows
    assert len(np.asarray(d.flat.obstacles).reshape(-1, 4)) == macros
    o = ol.Oracle.of(d)
    per = o.build_grid(bins, 1.0)
    rng = np.random.default_rng(seed)
    W, H = d.die[2], d.die[3]
    queries = []
    for k in range(200):
        x0, y0 = rng.uniform(0, W), rng.uniform(0, H)
        x1 = x0 + rng.uniform(0, W - x0) * (0.1 if k % 3 == 0 else 1.0)
        y1 = y0 + rng.uniform(0, H - y0) * (0.1 if k % 3 == 1 else 1.0)
        box, vertical = (x0, y0, x1, y1), bool(k % 2)
        queries.append(("P", box, vertical))
        # a half block's stripes: cut inside the box, the region's room at or
        # above the half's (zero: the half's own room), cells filling part of it
        lo, hi = (x0, x1) if vertical else (y0, y1)
        cut = lo + rng.uniform(0, 1) * (hi - lo)
        _, rp = o.profile(box, vertical)
        room = float(rp[-1]) if len(rp) else 0.0
        region = 0.0 if k % 5 == 0 else room * (1.0 + rng.uniform(0, 3))
        queries.append(("S", box, vertical, rng.uniform(0, 1.2) * room, region, 1.0, 0.1, cut, k % 4 < 2))
    for vertical in (True, False):
        queries.append(("P", (0.0, 0.0, W, H), vertical))
    avail, results = _run(harness, d, per, queries)
    ref_avail, _ = o.grid_avail()
    assert avail == _bits(ref_avail)
    stripes_seen = 0
    for q, got in zip(queries, results):
        if q[0] == "S":
            ref = o.stripes(q[1], q[2], q[3], q[4], q[7], q[8])
            assert got == _bits(ref.reshape(-1)), q
            stripes_seen += len(ref)
        else:
            re, rp = o.profile(q[1], q[2])
            assert got == (_bits(re), _bits(rp)), q
    assert stripes_seen > 400


def test_free_space_with_touching_and_overlapping_macros(harness):
    """FreeSpace::build on macros that overlap, touch, reach past the die or
    cover part of a row's height, and on rows split into touching or
    separated subrows: free area per bin and profiles equal the oracle's bit
    for bit."""
    cells = [(1.0, 0.0, 2.0, 1.0, False), (3.0, 0.0, 2.0, 1.0, False)]
    macros = [(10.0, 2.0, 20.0, 8.0), (15.0, 4.0, 25.0, 12.0),     # overlapping
              (30.0, 0.0, 34.0, 6.0), (34.0, 3.0, 38.0, 9.0),      # touching
              (55.0, 15.0, 70.0, 25.0),                            # past the die
              (40.5, 10.25, 45.5, 13.75),                          # part of four rows' height
              (2.0, 16.0, 3.0, 17.0), (2.0, 16.0, 3.0, 17.0)]      # the same twice
    cells += [(m[0], m[1], m[2] - m[0], m[3] - m[1], True) for m in macros]
    rows = [(float(r), 1.0, 0.0, 60.0) for r in range(20) if r not in (5, 7)]
    rows += [(5.0, 1.0, 0.0, 30.0), (5.0, 1.0, 30.0, 60.0), (7.0, 1.0, 0.0, 25.0), (7.0, 1.0, 35.0, 60.0)]
    rows.sort(key=lambda r: (r[0], r[2]))
    raw = dz.make_raw(cells, [[(0, 0.0, 0.0), (1, 0.0, 0.0)]], die=(0.0, 0.0, 60.0, 20.0), rows=rows)
    d = dz.collect(raw)
    o = ol.Oracle(raw)
    rng = np.random.default_rng(3)
    for bins in (1, 7, 16, 40):
        per = o.build_grid(bins, 1.0)
        queries = [("P", (0.0, 0.0, 60.0, 20.0), True), ("P", (0.0, 0.0, 60.0, 20.0), False)]
        for k in range(60):
            x0, x1 = np.sort(rng.uniform(-2, 62, 2))
            y0, y1 = np.sort(rng.uniform(-1, 21, 2))
            queries.append(("P", (x0, y0, x1, y1), bool(k % 2)))
        avail, results = _run(harness, d, per, queries)
        ref_avail, _ = o.grid_avail()
        assert avail == _bits(ref_avail), bins
        for q, got in zip(queries, results):
            re, rp = o.profile(q[1], q[2])
            assert got == (_bits(re), _bits(rp)), (bins, q)


@pytest.mark.parametrize("seed,bins", [(1, 16), (2, 37), (3, 64)])
def test_overfilled_clusters_match_label_and_oracle(harness, seed, bins):
    """overfilledClusters (a union-find labelling) against the oracle's flood
    fill and scipy.ndimage.label (4-connected), bounding boxes in the order
    of each cluster's first bin, on random occupancies with many touching
    and diagonal-only neighbours."""
    from scipy import ndimage
    d = dz.generate(3000, 3300, 60, seed=seed, num_macros=10)
    o = ol.Oracle.of(d)
    per = o.build_grid(bins, 1.0)
    avail, _ = o.grid_avail()
    rng = np.random.default_rng(seed)
    queries = []
    for frac in (0.2, 0.45, 0.6):
        occ = np.where(rng.random(per * per) < frac, avail * 1.5 + 1.0, avail * 0.5)
        queries.append(("C", 1.0, occ))
    _, results = _run(harness, d, per, queries)
    for (_, limit, occ), got in zip(queries, results):
        ref = o.overfilled_clusters(occ).reshape(-1).tolist()
        assert got == ref
        over = (occ > limit * avail).reshape(per, per)
        labels, count = ndimage.label(over, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
        want = [v for sl in ndimage.find_objects(labels)
                for v in (sl[1].start, sl[1].stop - 1, sl[0].start, sl[0].stop - 1)]
        assert got == want and count > 5


@pytest.mark.parametrize("seed,bins,macros,g", [(6, 16, 20, 1.0), (7, 0, 0, 0.8), (8, 48, 40, 0.9)])
def test_flat_geometry_of_the_batch_kernel_matches(harness, seed, bins, macros, g):
    """The fused GPU batch kernel computes profiles and stripes with the
    sequential steps of csrc/lal_seq.h on flat arrays (steps and cuts sorted
    as values, the stripe subdivision as a depth-first walk, the fill order
    by rank).  On the CPU those steps give the same profiles and stripes as
    lal_geom.cc and the oracle, bit for bit, on random boxes, thin boxes and
    boxes on row edges."""
    d = dz.generate(3000, 3300, 60, seed=seed, num_macros=macros)
    o = ol.Oracle.of(d)
    per = o.build_grid(bins, g)
    rng = np.random.default_rng(seed)
    W, H = d.die[2], d.die[3]
    rows = np.unique(np.asarray(d.raw.row_y, np.float64))
    queries = []
    for k in range(240):
        x0, y0 = rng.uniform(0, W), rng.uniform(0, H)
        if k % 4 == 3:  # on row edges
            y0 = float(rows[rng.integers(len(rows))])
        x1 = x0 + rng.uniform(0, W - x0) * (0.02 if k % 3 == 0 else 1.0)
        y1 = min(H, y0 + rng.uniform(0, H - y0) * (0.05 if k % 3 == 1 else 1.0))
        if k % 11 == 0:
            x1 = x0  # zero width
        box, vertical = (x0, y0, x1, y1), bool(k % 2)
        lo, hi = (x0, x1) if vertical else (y0, y1)
        cut = lo + rng.uniform(0, 1) * (hi - lo)
        _, rp = o.profile(box, vertical)
        room = float(rp[-1]) if len(rp) else 0.0
        region = 0.0 if k % 5 == 0 else room * (1.0 + rng.uniform(0, 3))
        frac = (0.1, 0.05, 0.3)[k % 3]
        s = (box, vertical, rng.uniform(0, 1.2) * room, region, g, frac, cut, k % 4 < 2)
        queries += [("P", box, vertical), ("p", box, vertical), ("S",) + s, ("s",) + s]
    _, results = _run(harness, d, per, queries)
    stripes_seen = 0
    for k in range(0, len(queries), 4):
        (qp, gp), (_, fp), (qs, gs), (_, fs) = zip(queries[k:k + 4], results[k:k + 4])
        assert fp == gp, qp
        assert fs == gs, qs
        ref = o.stripes(qs[1], qs[2], qs[3], qs[4], qs[7], qs[8]) if qs[5] == g and qs[6] == 0.1 else None
        if ref is not None:
            assert gs == _bits(ref.reshape(-1)), qs
        stripes_seen += len(fs) // 3
    assert stripes_seen > 300
