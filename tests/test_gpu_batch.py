"""Batched solve of many small designs against the two-system solver and the
oracle's systems.

Bar: every batched solution's true relative residual, computed on the CPU
with the oracle's own B2B matrix (scipy), at most the tolerance; iteration
counts within 2 of the two-system solver's on the same system (block sums
differ in order, so iterates agree to rounding) and its solution within 1e-6;
bit-identical results on repeat; the same solutions for every launch shape
and SpMV layout.
(The B2B systems are ill-conditioned, so solutions at a residual of 1e-10
can differ by far more than 1e-10: residuals are the invariant checked.)
"""
import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz
from ktplace_amd import simpl

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def instances():
    from ktplace_amd.batched import make_designs, start_context
    designs = make_designs(10, seed=9100)
    ctxs = [start_context(d, 1e-6) for d in designs]
    yield designs, ctxs
    for c in ctxs:
        c.close()


def test_batch_matches_oracle_and_pair_solver(instances):
    designs, ctxs = instances
    b = simpl.Batch(0)
    for c in ctxs:
        b.add_ctx(c)
    n, rows, nnz = b.size()
    assert n == len(ctxs) and rows == 2 * sum(d.n for d in designs)
    per, tot = b.solve(1e-10, 100000)
    for i, (d, c) in enumerate(zip(designs, ctxs)):
        x0, y0 = (v.copy() for v in c.get_positions())
        o = ol.Oracle.of(d)
        o.build_b2b(x0, y0, floors=d.flat.floors)
        gx, gy = b.get_positions(i)
        for ax, v in ((0, gx), (1, gy)):
            s = o.system(ax)
            true = np.linalg.norm(s.rhs - s.scipy() @ v) / np.linalg.norm(s.rhs)
            assert true <= 1.05e-10, (i, ax, true)
        c.save_positions()
        r = c.solve_resident(1e-10, 100000)
        px, py = c.get_positions()
        c.restore_positions()
        print(i, d.n, per[i]["steps"], r["steps"], flush=True)
        assert rel(gx, px) < 1e-6 and rel(gy, py) < 1e-6
        for ax in (0, 1):
            assert abs(per[i]["steps"][ax] - r["steps"][ax]) <= 2
            assert per[i]["resid"][ax] <= 1e-10
    b.close()


def test_batch_deterministic_and_shapes_agree(instances):
    _, ctxs = instances
    b = simpl.Batch(0)
    for c in ctxs[:6]:
        b.add_ctx(c)
    ref = None
    flat = {}
    for lanes, tile, rows, graphs, layout, rl, cb, wd in (
            (8, 1024, 2, True, 0, 0, 0, 0), (8, 1024, 2, True, 0, 0, 0, 0), (4, 256, 4, False, 0, 0, 0, 0),
            (16, 2048, 2, True, 0, 0, 0, 0), (16, 512, 4, True, 0, 0, 0, 0), (4, 512, 2, True, 1, 1, 0, 0),
            (4, 256, 2, False, 1, 4, 0, 0), (4, 1024, 2, True, 1, 4, 0, 0), (4, 1024, 2, True, 1, 4, 3, 0),
            (4, 1024, 2, False, 1, 4, 3, 0), (4, 1024, 2, True, 1, 4, 3, 1), (4, 1024, 2, False, 1, 4, 4, 1)):
        b.set_tuning(lanes=lanes, tile=tile, rows=rows, use_graphs=graphs, layout=layout, row_lanes=rl,
                     col_bytes=cb, wide=wd)
        b.reset()
        per, _ = b.solve(1e-8, 100000, simpl.KT_STOP_TRUE)
        pos = [b.get_positions(i) for i in range(6)]
        assert all(max(p["true_resid"]) <= 1e-8 for p in per)
        if (tile, layout, rl) == (1024, 1, 4):  # 3- and 4-byte columns, wide loads: bit-identical
            if flat:
                for (ax, ay), (bx, by) in zip(pos, flat["pos"]):
                    assert np.array_equal(ax, bx) and np.array_equal(ay, by)
            flat["pos"] = pos
        if ref is None:
            ref = (per, pos)
            continue
        if (lanes, tile, rows, layout) == (8, 1024, 2, 0):  # same shape: bit-identical
            assert [p["steps"] for p in per] == [p["steps"] for p in ref[0]]
            for (ax, ay), (bx, by) in zip(pos, ref[1]):
                assert np.array_equal(ax, bx) and np.array_equal(ay, by)
        else:
            for (ax, ay), (bx, by) in zip(pos, ref[1]):
                assert rel(ax, bx) < 1e-6 and rel(ay, by) < 1e-6
    b.close()


@pytest.mark.parametrize("layout,wide", [(0, 0), (1, 0), (1, 1)])
def test_batch_edge_cases(layout, wide):
    """A zero right-hand side (axis inactive), a single-row system, a row with
    no off-diagonals and the max_iter cap, next to an ordinary instance, for
    both SpMV layouts and the entry blocks' wide loads."""
    from test_oracle import spd_system
    rng = np.random.default_rng(3)
    b = simpl.Batch(0)
    b.set_tuning(layout=layout, row_lanes=4 if layout else 0, col_bytes=3 if wide else 0, wide=wide)
    s1, s2 = spd_system(rng, 500, 3000), spd_system(rng, 500, 2000)
    zero = simpl.SparseSystem(s2.row_ptr, s2.col, s2.val, s2.diag, np.zeros(500))
    b.add(s1, zero, np.zeros(500), np.zeros(500))
    one = simpl.SparseSystem(np.array([0, 0], np.uint64), np.zeros(0, np.uint32), np.zeros(0),
                             np.array([2.0]), np.array([4.0]))
    b.add(one, one, np.zeros(1), np.ones(1))
    per, tot = b.solve(1e-12, 100000)
    assert per[0]["steps"][1] == 0 and per[0]["resid"][0] <= 1e-12
    x, y = b.get_positions(0)
    assert np.array_equal(y, np.zeros(500))
    from scipy.sparse.linalg import spsolve
    assert rel(x, spsolve(s1.scipy().tocsc(), s1.rhs)) < 1e-9
    x1, y1 = b.get_positions(1)
    assert x1[0] == pytest.approx(2.0) and y1[0] == pytest.approx(2.0)
    b.reset()
    per, _ = b.solve(1e-12, 5)
    assert per[0]["steps"][0] == 5
    b.close()


def test_batch_all_256_instances_of_the_benchmark():
    """The benchmark's batched workload at its full size: 256 netlists of
    2,000-40,000 cells (log-uniform), every x and y system solved in one
    batch to a true relative residual of 1e-6, each solution's residual
    checked on the CPU with the oracle's own B2B matrix."""
    from ktplace_amd.batched import make_designs, start_context
    designs = make_designs(256)
    ctxs = []
    b = simpl.Batch(0)
    try:
        for k, d in enumerate(designs):
            ctxs.append(start_context(d, 1e-6))
            b.add_ctx(ctxs[-1])
            if k % 64 == 63:
                print(f"{k + 1} instances ready", flush=True)
        n, rows, nnz = b.size()
        assert n == 256 and rows == 2 * sum(d.n for d in designs)
        per, tot = b.solve(1e-6, 100000, simpl.KT_STOP_TRUE)
        print(f"batched: {rows} rows, {nnz} entries, {tot}", flush=True)
        worst = 0.0
        for i, (d, c) in enumerate(zip(designs, ctxs)):
            x0, y0 = c.get_positions()  # the start point the instance was added with
            o = ol.Oracle.of(d)
            o.build_b2b(x0, y0, floors=d.flat.floors)
            gx, gy = b.get_positions(i)
            for ax, v in ((0, gx), (1, gy)):
                s = o.system(ax)
                true = np.linalg.norm(s.rhs - s.scipy() @ v) / np.linalg.norm(s.rhs)
                worst = max(worst, true)
                assert true <= 1.05e-6, (i, d.n, ax, true)
            assert max(per[i]["true_resid"]) <= 1e-6
        print(f"worst true residual over 512 systems (oracle matrices): {worst:.3e}", flush=True)
    finally:
        b.close()
        for c in ctxs:
            c.close()
