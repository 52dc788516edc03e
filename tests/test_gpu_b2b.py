"""B2B assembly on the GPU against the oracle: the CSR must match bit for bit."""
import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz

pytestmark = pytest.mark.gpu


def assert_same_system(gpu, ora, label):
    assert np.array_equal(gpu.row_ptr, ora.row_ptr), f"{label}: rowPtr differs"
    assert np.array_equal(gpu.col, ora.col), f"{label}: col differs"
    # bit-exact doubles (view as integers so -0.0 / NaN patterns count too)
    for name in ("val", "diag", "rhs"):
        g = getattr(gpu, name).view(np.uint64)
        o = getattr(ora, name).view(np.uint64)
        bad = np.flatnonzero(g != o)
        assert len(bad) == 0, f"{label}: {name} differs at {bad[:5]} ({getattr(gpu, name)[bad[:3]]} vs {getattr(ora, name)[bad[:3]]})"


@pytest.fixture(scope="module", params=["cells12k", "cells210k"])
def case(request):
    d = dz.workload(request.param)
    return d, ol.Oracle.of(d)


@pytest.mark.parametrize("alpha", [0.0, 0.06])
def test_b2b_bit_exact_random_positions(case, gpu_ctx_factory, alpha):
    d, o = case
    rng = np.random.default_rng(17)
    x = rng.uniform(0, d.die[2] - 8, d.n)
    y = rng.uniform(0, d.die[3] - 1, d.n)
    ax = rng.uniform(0, d.die[2] - 8, d.n) if alpha else None
    ay = rng.uniform(0, d.die[3] - 1, d.n) if alpha else None
    print(f"b2b {d.name} n={d.n} alpha={alpha}", flush=True)
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x, y, alpha, ax, ay)
    o.build_b2b(x, y, alpha, ax, ay, floors=d.flat.floors)
    for axis in (0, 1):
        assert_same_system(ctx.get_system(axis), o.system(axis), f"{d.name} axis {axis}")
    ctx.close()


def test_b2b_bit_exact_ties_and_collapsed_nets(gpu_ctx_factory):
    """All cells at a handful of positions: many equal pins (first-index tie
    break) and nets whose pins coincide (lo == hi, skipped)."""
    raw = dz.generate(5000, 5600, 100, seed=5).raw
    raw.pin_off_x[:] = np.round(raw.pin_off_x)  # coinciding pins are common
    d = dz.collect(raw, "ties")
    o = ol.Oracle(raw)
    rng = np.random.default_rng(3)
    x = rng.integers(0, 4, d.n).astype(float)
    y = rng.integers(0, 3, d.n).astype(float)
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x, y)
    o.build_b2b(x, y, floors=d.flat.floors)
    for axis in (0, 1):
        assert_same_system(ctx.get_system(axis), o.system(axis), f"ties axis {axis}")
    ctx.close()


def test_b2b_bit_exact_one_million_cells(gpu_ctx_factory):
    d = dz.workload("cells1m")
    o = ol.Oracle.of(d)
    x, y = o.seed_uniform(20240607)
    print(f"b2b cells1m n={d.n}", flush=True)
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x, y)
    o.build_b2b(x, y, floors=d.flat.floors)
    for axis in (0, 1):
        g = ctx.get_system(axis)
        assert_same_system(g, o.system(axis), f"cells1m axis {axis}")
    # size-independent properties at full size: symmetric pattern, positive weights
    rows = np.repeat(np.arange(d.n), np.diff(g.row_ptr.astype(np.int64)))
    assert (g.val < 0).all()
    key_f = np.sort(rows.astype(np.int64) * d.n + g.col)
    key_b = np.sort(g.col.astype(np.int64) * d.n + rows)
    assert np.array_equal(key_f, key_b)
    ctx.close()


def test_b2b_rebuild_reuses_buffers(gpu_ctx_factory):
    """Consecutive builds on one context (the SimPL loop's pattern) stay exact."""
    d = dz.workload("cells12k")
    o = ol.Oracle.of(d)
    ctx = gpu_ctx_factory(d)
    rng = np.random.default_rng(9)
    for it in range(3):
        x = rng.uniform(0, d.die[2], d.n)
        y = rng.uniform(0, d.die[3], d.n)
        alpha = 0.02 * it
        ctx.build_b2b(x, y, alpha, x[::-1].copy(), y[::-1].copy())
        o.build_b2b(x, y, alpha, x[::-1].copy(), y[::-1].copy(), floors=d.flat.floors)
        for axis in (0, 1):
            assert_same_system(ctx.get_system(axis), o.system(axis), f"rebuild {it} axis {axis}")
    ctx.close()


def test_multiply_bit_exact(gpu_ctx_factory):
    d = dz.workload("cells12k")
    o = ol.Oracle.of(d)
    x, y = o.seed_uniform(1)
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x, y)
    o.build_b2b(x, y, floors=d.flat.floors)
    v = np.random.default_rng(2).standard_normal(d.n)
    for axis in (0, 1):
        assert np.array_equal(ctx.multiply(axis, v), o.multiply(axis, v))
    ctx.close()


@pytest.mark.parametrize("name", ["cells12k", "cells210k"])
def test_star_bit_exact(gpu_ctx_factory, name):
    """LinearSystems::buildStar on the GPU: same variables, CSR, diag and rhs
    as the oracle, bit for bit."""
    d = dz.workload(name)
    o = ol.Oracle.of(d)
    ctx = gpu_ctx_factory(d)
    size = ctx.build_star()
    assert size == o.build_star()
    assert size > d.n
    for axis in (0, 1):
        g, r = ctx.get_system(axis), o.system(axis)
        assert np.array_equal(g.row_ptr, r.row_ptr)
        assert np.array_equal(g.col, r.col)
        assert np.array_equal(g.val, r.val)
        assert np.array_equal(g.diag, r.diag)
        assert np.array_equal(g.rhs, r.rhs)
    ctx.close()
