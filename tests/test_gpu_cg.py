"""Jacobi-PCG on the GPU (FP64) against the oracle and scipy.

Tolerances: the GPU reduces dot products in a different (fixed) order than
the reference's 4096-element chunks, so iterates agree to rounding, not bit
for bit.  Solutions are compared with a relative 2-norm error bound of 1e-6
against the oracle solved to 1e-12.
"""
import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz
from test_oracle import line_design, spd_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def set_both(ctx, o, sx, sy):
    for axis, s in ((0, sx), (1, sy)):
        ctx.set_system(axis, s)
        o.set_system(axis, s)


def test_cg_random_spd_vs_scipy(gpu_ctx_factory):
    from scipy.sparse.linalg import spsolve
    rng = np.random.default_rng(21)
    sx, sy = spd_system(rng, 3000, 12000), spd_system(rng, 3000, 9000)
    ctx = gpu_ctx_factory()
    ctx.set_system(0, sx)
    ctx.set_system(1, sy)
    x = np.zeros(sx.n)
    y = np.zeros(sy.n)
    r = ctx.solve_both(x, y, 1e-12, 100000)
    for sol, s in ((x, sx), (y, sy)):
        exact = spsolve(s.scipy().tocsc(), s.rhs)
        assert rel(sol, exact) < 1e-9
    assert max(r["true_resid"]) < 1e-11
    ctx.close()


@pytest.fixture(scope="module")
def b2b12k():
    d = dz.workload("cells12k")
    o = ol.Oracle.of(d)
    x0, y0 = o.seed_uniform(20240607)
    o.build_b2b(x0, y0, floors=d.flat.floors)
    return d, o, x0, y0


def test_cg_tracks_oracle_iterates(b2b12k, gpu_ctx_factory):
    """Same start, same tolerance: same iteration count (+-1) and solutions
    within 1e-6; the per-iteration residuals (observer) agree closely."""
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    seen = []
    r = ctx.solve_both(x, y, 1e-6, 100000, observe=lambda it, res, last: seen.append((it, res, last)))
    ref = o.solve(x0, y0, tol=1e-6, max_iter=100000, history=True)
    print("gpu iters", r["iterations"], "oracle iters", ref["iterations"], flush=True)
    assert abs(r["iterations"] - ref["iterations"]) <= 1
    assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
    assert [s[0] for s in seen] == list(range(1, r["iterations"] + 1))
    assert seen[-1][2] and not any(s[2] for s in seen[:-1])
    hist = np.array([s[1] for s in seen])
    k = min(len(hist), len(ref["history"])) - 2
    assert np.allclose(hist[:k], ref["history"][:k], rtol=1e-6)
    assert max(r["resid"]) <= 1e-6


def test_cg_solution_within_1e6_of_tight_oracle(b2b12k, gpu_ctx_factory):
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    r = ctx.solve_both(x, y, 1e-9, 100000)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
    assert max(r["true_resid"]) < 1e-8


def test_cg_rules_zero_rhs_maxiter_breakdown(gpu_ctx_factory):
    rng = np.random.default_rng(5)
    sx = spd_system(rng)
    zero = ol.System(sx.row_ptr, sx.col, sx.val, sx.diag, np.zeros(sx.n))
    neg = ol.System(sx.row_ptr, sx.col, -sx.val, -sx.diag, sx.rhs)
    o = ol.Oracle(line_design([1.0, 2.0]))
    ctx = gpu_ctx_factory()
    x0 = np.zeros(sx.n)
    # zero rhs: axis inactive, residual 0
    set_both(ctx, o, sx, zero)
    x, y = x0.copy(), x0.copy()
    r = ctx.solve_both(x, y, 1e-8, 5000)
    ref = o.solve(x0, x0, tol=1e-8, max_iter=5000)
    assert r["resid"][1] == 0.0 and r["steps"][1] == 0
    assert abs(r["iterations"] - ref["iterations"]) <= 1
    assert np.array_equal(y, x0)
    # max_iter = 0
    x, y = x0.copy(), x0.copy()
    r = ctx.solve_both(x, y, 1e-8, 0)
    assert r["iterations"] == 0 and np.array_equal(x, x0)
    # indefinite: p.Ap <= 0 stops the y axis after one step, x continues
    set_both(ctx, o, sx, neg)
    x, y = x0.copy(), x0.copy()
    seen = []
    r = ctx.solve_both(x, y, 1e-8, 50, observe=lambda it, res, last: seen.append(res))
    ref = o.solve(x0, x0, tol=1e-8, max_iter=50, history=True)
    assert r["steps"][1] == 1 and r["iterations"] == ref["iterations"]
    assert np.array_equal(y, x0)
    assert all(s[1] == seen[0][1] for s in seen)
    ctx.close()


def test_cg_max_iter_cap_and_warm_start(b2b12k, gpu_ctx_factory):
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    r = ctx.solve_both(x, y, 1e-6, 60)  # the reference's hidden cap
    ref = o.solve(x0, y0, tol=1e-6, max_iter=60)
    assert r["iterations"] == ref["iterations"] == 60
    assert rel(x, ref["x"]) < 1e-9
    # warm start: continuing from the capped iterate converges in fewer steps
    r2 = ctx.solve_both(x, y, 1e-6, 100000)
    ref2 = o.solve(ref["x"], ref["y"], tol=1e-6, max_iter=100000)
    assert abs(r2["iterations"] - ref2["iterations"]) <= 1
    ctx.close()


def test_cg_isolated_cell_floor(gpu_ctx_factory):
    """A row with no entries and diag 1e-12 (isolated cell) solves like the oracle."""
    rng = np.random.default_rng(8)
    s = spd_system(rng, 200, 500)
    rp = s.row_ptr.astype(np.int64)
    # empty row 7: drop its entries and the mirrored ones
    keep = np.ones(len(s.col), bool)
    rows = np.repeat(np.arange(s.n), np.diff(rp))
    keep &= ~((rows == 7) | (s.col == 7))
    rows, col, val = rows[keep], s.col[keep], s.val[keep]
    diag = s.diag.copy()
    diag[7] = 1e-12
    rhs = s.rhs.copy()
    rhs[7] = 0.0
    new_rp = np.zeros(s.n + 1, np.uint64)
    new_rp[1:] = np.cumsum(np.bincount(rows, minlength=s.n))
    s2 = ol.System(new_rp, col, val, diag, rhs)
    o = ol.Oracle(line_design([1.0, 2.0]))
    ctx = gpu_ctx_factory()
    set_both(ctx, o, s2, s2)
    x0 = np.zeros(s.n)
    x, y = x0.copy(), x0.copy()
    r = ctx.solve_both(x, y, 1e-10, 10000)
    ref = o.solve(x0, x0, tol=1e-10, max_iter=10000)
    assert rel(x, ref["x"]) < 1e-8 and x[7] == 0.0
    ctx.close()


# ---------------------------------------------------------------- mixed precision
def test_mixed_precision_reaches_true_tolerance(b2b12k, gpu_ctx_factory):
    """FP32 PCG with FP64 refreshes: the true FP64 residual reaches tol and the
    solution agrees with the tightly solved FP64 oracle."""
    from ktplace_amd.simpl import KT_PREC_MIXED
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    for tol in (1e-6, 1e-9):
        x, y = x0.copy(), y0.copy()
        r = ctx.solve_both(x, y, tol, 100000, precision=KT_PREC_MIXED)
        print("mixed", tol, r, flush=True)
        assert max(r["true_resid"]) <= tol
        assert r["refinements"] >= 2
        if tol <= 1e-9:
            assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
        # independent check of the reported residual with the oracle's SpMV
        for axis, sol in ((0, x), (1, y)):
            s = o.system(axis)
            true = np.linalg.norm(s.rhs - o.multiply(axis, sol)) / np.linalg.norm(s.rhs)
            assert true <= tol * 1.0001
    ctx.close()


@pytest.mark.parametrize("tol", [1e-6, 1e-9])
def test_fused_update_matches_three_kernel_iteration(b2b12k, gpu_ctx_factory, tol):
    """The two-kernel FP32 iteration (the SpMV also sums r.Ap and Ap.Ap; one
    kernel updates r, the correction and p with r'.r' = r.r - 2 alpha r.Ap +
    alpha^2 Ap.Ap) against the three-kernel iteration on the same scaled
    system: both reach the true FP64 tolerance, agree with the oracle solved
    to 1e-12 (at 1e-9), step the same number of times give or take 3 %, a
    profiled solve (one launch per kernel) repeats the graph-replayed one bit
    for bit, and the fused solve is the one that ran: its profile has no
    direction update kernel.  A fused step with a wrong expansion of r'.r'
    (a sign or a factor of alpha) gives a wrong beta and residual, and the
    solve misses the tolerance or the count."""
    from ktplace_amd.simpl import (KERNEL_PUPDATE, KERNEL_UPDATE, KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS,
                                   KT_STOP_TRUE)
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    ctx.set_tuning(lanes=4, tile=1024, rows=2, layout=KT_SPMV_ENTRY_BLOCKS, block_entries=2048, unroll=4,
                   row_lanes=4, block_threads=256, col_bytes=3, wide=1)
    runs = {}
    for on in (True, False):
        ctx.set_mixed_fused_update(on)
        for profile in (False, True):
            ctx.reset_kernel_times()
            x, y = x0.copy(), y0.copy()
            r = ctx.solve_both(x, y, tol, 100000, precision=KT_PREC_MIXED, stop_rule=KT_STOP_TRUE,
                               profile=profile)
            print("fused" if on else "three kernels", "profiled" if profile else "graphs",
                  r["steps"], r["refinements"], r["true_resid"], flush=True)
            assert max(r["true_resid"]) <= tol
            if tol <= 1e-9:
                assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
            if profile:
                updates, directions = ctx.kernel_time(KERNEL_UPDATE)[0], ctx.kernel_time(KERNEL_PUPDATE)[0]
                assert updates > 0 and (directions == 0 if on else directions > 0)
                assert np.array_equal(x, runs[on][0]) and np.array_equal(y, runs[on][1])
            else:
                runs[on] = (x, y, r)
    for ax in (0, 1):
        a, b = runs[True][2]["steps"][ax], runs[False][2]["steps"][ax]
        assert abs(a - b) <= max(2, 0.03 * b)
    ctx.close()


@pytest.mark.parametrize("layout", ["lane groups", "entry blocks", "entry blocks, wide"])
def test_mixed_scaled_and_unscaled_agree(b2b12k, gpu_ctx_factory, layout):
    """The FP32 iterations on D^-1/2 A D^-1/2 (the default) and on A: both
    reach a true FP64 residual of 1e-9 within the same iteration count give
    or take 3 %, agree with the oracle solved to 1e-12, and the scaled solve
    repeats bit for bit.  A scaled solve that left out a D^-1/2 anywhere
    (start, refresh or correction) would stall or miss the oracle."""
    from ktplace_amd.simpl import KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS, KT_STOP_TRUE
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    if layout == "lane groups":
        tuning = dict(lanes=4, tile=1024, rows=2)
    else:
        tuning = dict(lanes=4, tile=1024, rows=2, layout=KT_SPMV_ENTRY_BLOCKS, block_entries=2048, unroll=4,
                      row_lanes=4, block_threads=256, col_bytes=3, wide=int(layout.endswith("wide")))
    ctx.set_tuning(**tuning)
    runs = {}
    for scaled in (True, False, True):
        ctx.set_mixed_scaling(scaled)
        x, y = x0.copy(), y0.copy()
        r = ctx.solve_both(x, y, 1e-9, 100000, precision=KT_PREC_MIXED, stop_rule=KT_STOP_TRUE)
        print(layout, "scaled" if scaled else "unscaled", r["steps"], r["refinements"], r["true_resid"], flush=True)
        assert max(r["true_resid"]) <= 1e-9
        assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
        if scaled in runs:
            assert np.array_equal(runs[scaled][0], x) and np.array_equal(runs[scaled][1], y)
        runs[scaled] = (x, y, r)
    for ax in (0, 1):
        a, b = runs[True][2]["steps"][ax], runs[False][2]["steps"][ax]
        assert abs(a - b) <= max(2, 0.03 * b)
    ctx.close()


@pytest.mark.parametrize("lanes,tile,graphs,rows", [(1, 1024, False, 2), (2, 2048, True, 4), (8, 4096, True, 8),
                                                    (16, 8192, False, 2)])
def test_launch_shapes_agree(b2b12k, gpu_ctx_factory, lanes, tile, graphs, rows):
    """Every tuning point solves the same system to the same answer."""
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.set_tuning(lanes=lanes, tile=tile, use_graphs=graphs, rows=rows)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    r = ctx.solve_both(x, y, 1e-9, 100000)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
    assert max(r["true_resid"]) < 1e-8
    ctx.close()


def test_solve_is_deterministic(b2b12k, gpu_ctx_factory):
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    outs = []
    for _ in range(2):
        x, y = x0.copy(), y0.copy()
        ctx.solve_both(x, y, 1e-7, 100000)
        outs.append((x, y))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    ctx.close()


@pytest.mark.parametrize("xcd", [0, 1])
def test_xcd_split_agrees(b2b12k, gpu_ctx_factory, xcd):
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.set_tuning(lanes=4, tile=2048, xcd_split=xcd)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    ctx.solve_both(x, y, 1e-9, 100000)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
    ctx.close()


def test_exact_mode_is_bit_identical_to_oracle(b2b12k, gpu_ctx_factory):
    """KT_PREC_FP64_EXACT reproduces the reference's operation order."""
    from ktplace_amd.simpl import KT_PREC_FP64_EXACT
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    x, y = x0.copy(), y0.copy()
    seen = []
    r = ctx.solve_both(x, y, 1e-6, 60, precision=KT_PREC_FP64_EXACT,
                       observe=lambda it, res, last: seen.append(res))
    ref = o.solve(x0, y0, tol=1e-6, max_iter=60, history=True)
    assert r["iterations"] == ref["iterations"]
    assert np.array_equal(x, ref["x"]) and np.array_equal(y, ref["y"])
    assert np.array_equal(np.array(seen), ref["history"])
    ctx.close()


@pytest.mark.parametrize("precision,lanes,tile,halo,xcd", [
    ("fp64", 8, 2048, 256, 1), ("fp64", 4, 1024, 4096, 0), ("fp64", 16, 1024, 8192, 1),
    ("mixed", 16, 2048, 7168, 1), ("mixed", 4, 1024, 16384, 0), ("mixed", 8, 4096, 2048, 1)])
def test_lds_window_is_bit_identical(b2b12k, gpu_ctx_factory, precision, lanes, tile, halo, xcd):
    """Staging p in LDS changes where the SpMV reads p, not what it reads:
    iterates, residual history and solutions equal the unstaged solve's bit
    for bit (windows clipped at both ends of the vector, and one wider than
    the whole system)."""
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED
    prec = KT_PREC_FP64 if precision == "fp64" else KT_PREC_MIXED
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    outs = []
    for h in (0, halo):
        ctx.set_tuning(lanes=lanes, tile=tile, xcd_split=xcd, halo=h)
        x, y = x0.copy(), y0.copy()
        seen = []
        r = ctx.solve_both(x, y, 1e-8, 100000, precision=prec, observe=lambda it, res, last: seen.append(res))
        outs.append((x, y, np.array(seen), r["steps"]))
    (x1, y1, h1, s1), (x2, y2, h2, s2) = outs
    assert s1 == s2
    assert np.array_equal(h1, h2)
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    ctx.close()


def test_lds_window_ragged_tail_and_limit(gpu_ctx_factory):
    """A system whose row count is not a multiple of the 16-byte staging
    width, and a window beyond the LDS budget (rejected, not truncated)."""
    from ktplace_amd.simpl import KT_PREC_MIXED, KtError
    rng = np.random.default_rng(5)
    sx, sy = spd_system(rng, 3001, 12000), spd_system(rng, 3001, 9000)
    ctx = gpu_ctx_factory()
    ctx.set_system(0, sx)
    ctx.set_system(1, sy)
    outs = []
    for halo in (0, 512):
        for prec in (0, KT_PREC_MIXED):
            ctx.set_tuning(lanes=4, tile=1024, halo=halo)
            x, y = np.zeros(sx.n), np.zeros(sy.n)
            ctx.solve_both(x, y, 1e-10, 100000, precision=prec)
            outs.append((x, y))
    for a, b in ((outs[0], outs[2]), (outs[1], outs[3])):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ctx.set_tuning(lanes=4, tile=2048, halo=16384)  # (2048 + 32768) * 8 bytes
    with pytest.raises(KtError):
        ctx.solve_both(np.zeros(sx.n), np.zeros(sy.n), 1e-10, 100, precision=0)
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tolerance_below_fp32_accuracy_on_ragged_system(gpu_ctx_factory, precision):
    """A well-conditioned system with an empty row, solved to 1e-10, below
    what FP32 arithmetic resolves: the mixed path must get there by refreshes (an axis pauses at each refresh point;
    FP32 steps run past it once drove p far below the refreshed residual and
    the next step diverged), match scipy, and repeat bit for bit."""
    from scipy.sparse.linalg import spsolve
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED, KT_STOP_TRUE
    prec = KT_PREC_FP64 if precision == "fp64" else KT_PREC_MIXED
    rng = np.random.default_rng(11)
    base = spd_system(rng, 2501, 11000)
    d = base.scipy().tolil()
    d[5, :] = 0
    d[:, 5] = 0
    d = d.tocsr()
    d.setdiag(base.diag + np.abs(d).sum(axis=1).A1 + 1.0)
    off = d.copy()
    off.setdiag(0)
    off.eliminate_zeros()
    off.sort_indices()
    sx = ol.System(off.indptr.astype(np.uint64), off.indices.astype(np.uint32), off.data.copy(),
                   d.diagonal().copy(), rng.standard_normal(2501))
    assert sx.row_ptr[6] == sx.row_ptr[5]
    ctx = gpu_ctx_factory()
    ctx.set_system(0, sx)
    ctx.set_system(1, base)
    outs = []
    for _ in range(2):
        ctx.set_tuning(lanes=2, tile=1024, rows=4)
        x, y = np.zeros(2501), np.zeros(2501)
        r = ctx.solve_both(x, y, 1e-10, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
        outs.append((x, y))
        assert max(r["true_resid"]) <= 1e-10
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    for sol, s in ((outs[0][0], sx), (outs[0][1], base)):
        assert rel(sol, spsolve(s.scipy().tocsc(), s.rhs)) < 1e-8
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_entry_block_layout_solves_and_repeats(b2b12k, gpu_ctx_factory, precision):
    """The entry-block SpMV: same answer as the oracle solved tight, iteration
    count within 1 of the lane-group layout's, bit-identical on repeat."""
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS, KT_STOP_TRUE
    prec = KT_PREC_FP64 if precision == "fp64" else KT_PREC_MIXED
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.build_b2b(x0, y0)
    ref = o.solve(x0, y0, tol=1e-12, max_iter=100000)
    runs = []
    cases = ((0, 0, 0), (KT_SPMV_ENTRY_BLOCKS, 4, 0), (KT_SPMV_ENTRY_BLOCKS, 4, 0), (KT_SPMV_ENTRY_BLOCKS, 1, 0),
             (KT_SPMV_ENTRY_BLOCKS, 4, 256), (KT_SPMV_ENTRY_BLOCKS, 4, 256), (KT_SPMV_ENTRY_BLOCKS, 1, 512))
    for layout, row_lanes, threads in cases:
        ctx.set_tuning(lanes=4, tile=2048, layout=layout, row_lanes=row_lanes, block_threads=threads)
        x, y = x0.copy(), y0.copy()
        r = ctx.solve_both(x, y, 1e-9, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
        runs.append((x, y, r))
    for x, y, r in runs[1:]:
        assert rel(x, ref["x"]) < 1e-6 and rel(y, ref["y"]) < 1e-6
        assert max(r["true_resid"]) <= 1e-9
        assert all(abs(a - b) <= 1 for a, b in zip(r["steps"], runs[0][2]["steps"]))
    for a, b in ((1, 2), (4, 5)):
        assert np.array_equal(runs[a][0], runs[b][0]) and np.array_equal(runs[a][1], runs[b][1])
    # a change of row lanes or block threads alone must not replay the graph
    # captured for the previous shape: graph replay equals direct launches
    for row_lanes, threads in ((4, 256), (1, 256), (1, 1024), (4, 1024)):
        out = []
        for graphs in (True, False):
            ctx.set_tuning(lanes=4, tile=2048, layout=KT_SPMV_ENTRY_BLOCKS, row_lanes=row_lanes,
                           block_threads=threads, use_graphs=graphs)
            x, y = x0.copy(), y0.copy()
            ctx.solve_both(x, y, 1e-9, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
            out.append((x, y))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    ctx.close()


def test_entry_block_layout_long_rows(gpu_ctx_factory):
    """Rows longer than an entry block's LDS share (a hub row of 12,000 entries
    and one of 9,000 crossing the block cuts, done by the lane-group path), empty
    rows and a last block of a few entries, for each block size, thread count
    and unroll: the solution matches scipy."""
    from scipy.sparse.linalg import spsolve
    import scipy.sparse as sp
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS, KT_STOP_TRUE
    rng = np.random.default_rng(17)
    n = 30001
    i = np.concatenate([np.zeros(12000, int), np.full(9000, 20000), rng.integers(0, n, 40000)])
    j = np.concatenate([rng.choice(np.arange(1, n), 12000, replace=False),
                        rng.choice(np.arange(0, 20000), 9000, replace=False), rng.integers(0, n, 40000)])
    keep = (i != j) & (i != 7) & (j != 7)          # row 7 stays empty
    i, j = i[keep], j[keep]
    w = rng.uniform(0.1, 2.0, len(i))
    off = sp.coo_matrix((-np.concatenate([w, w]), (np.concatenate([i, j]), np.concatenate([j, i]))),
                        shape=(n, n)).tocsr()
    off.sum_duplicates()
    off.sort_indices()
    diag = -off.sum(axis=1).A1 + rng.uniform(0.05, 0.5, n)
    s = ol.System(off.indptr.astype(np.uint64), off.indices.astype(np.uint32), off.data.copy(), diag,
                  rng.standard_normal(n))
    lengths = np.diff(off.indptr)
    assert lengths.max() >= 12000 and lengths[7] == 0
    exact = spsolve(s.scipy().tocsc(), s.rhs)
    ctx = gpu_ctx_factory()
    ctx.set_system(0, s)
    ctx.set_system(1, s)
    for prec, entries, unroll, row_lanes, threads in (
            (KT_PREC_FP64, 0, 0, 0, 0), (KT_PREC_MIXED, 0, 0, 4, 0), (KT_PREC_FP64, 4096, 8, 4, 0),
            (KT_PREC_MIXED, 16384, 8, 1, 0), (KT_PREC_FP64, 16384, 4, 4, 0), (KT_PREC_FP64, 2048, 4, 4, 256),
            (KT_PREC_MIXED, 0, 8, 1, 256), (KT_PREC_MIXED, 4096, 4, 4, 512), (KT_PREC_FP64, 0, 0, 1, 512),
            (KT_PREC_FP64, 2048, -4, 4, 256), (KT_PREC_MIXED, 2048, -8, 4, 256)):
        # unroll < 0: wide loads with |unroll| entries per thread and round
        ctx.set_tuning(lanes=4, tile=1024, layout=KT_SPMV_ENTRY_BLOCKS, block_entries=entries, unroll=abs(unroll),
                       row_lanes=row_lanes, block_threads=threads, wide=int(unroll < 0))
        x, y = np.zeros(n), np.zeros(n)
        r = ctx.solve_both(x, y, 1e-10, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
        assert max(r["true_resid"]) <= 1e-10
        assert rel(x, exact) < 1e-8 and np.array_equal(x, y)
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_three_byte_columns_are_bit_identical(b2b12k, gpu_ctx_factory, precision):
    """Entry blocks reading column indices as 16 + 8 bits give the same
    solution bit for bit as reading them as 32 bits (same products, same
    order), through graph replay and direct launches, after the systems are
    rebuilt (the 3-byte copy follows the new columns)."""
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS, KT_STOP_TRUE
    prec = KT_PREC_FP64 if precision == "fp64" else KT_PREC_MIXED
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    for positions in ((x0, y0), (x0[::-1].copy(), y0)):
        ctx.build_b2b(*positions)
        out = []
        for col_bytes, graphs in ((4, True), (3, True), (3, False)):
            ctx.set_tuning(lanes=4, tile=2048, layout=KT_SPMV_ENTRY_BLOCKS, block_threads=256, row_lanes=4,
                           col_bytes=col_bytes, use_graphs=graphs)
            x, y = x0.copy(), y0.copy()
            r = ctx.solve_both(x, y, 1e-9, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
            assert max(r["true_resid"]) <= 1e-9
            out.append((x, y))
        for x, y in out[1:]:
            assert np.array_equal(x, out[0][0]) and np.array_equal(y, out[0][1])
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_loads_are_bit_identical(b2b12k, gpu_ctx_factory, precision):
    """Entry blocks reading four consecutive entries per thread by wide loads
    (from each block's first entry rounded down to a multiple of four, with 3-
    or 4-byte columns, 4 or 8 entries per round) give the same solution bit for
    bit as one entry per load, through graph replay and direct launches, after
    the systems are rebuilt.  Mixed precision runs the three-kernel iteration
    here (the fused one exists only in the wide 3-byte shape; it has its own
    test)."""
    from ktplace_amd.simpl import KT_PREC_FP64, KT_PREC_MIXED, KT_SPMV_ENTRY_BLOCKS, KT_STOP_TRUE
    prec = KT_PREC_FP64 if precision == "fp64" else KT_PREC_MIXED
    d, o, x0, y0 = b2b12k
    ctx = gpu_ctx_factory(d)
    ctx.set_mixed_fused_update(False)
    for positions in ((x0, y0), (x0[::-1].copy(), y0)):
        ctx.build_b2b(*positions)
        out = []
        for wide, unroll, col_bytes, graphs in ((0, 4, 4, True), (1, 4, 3, True), (1, 8, 3, True), (1, 4, 4, True),
                                               (1, 8, 4, False), (1, 4, 3, False)):
            ctx.set_tuning(lanes=4, tile=2048, layout=KT_SPMV_ENTRY_BLOCKS, block_threads=256, row_lanes=4,
                           unroll=unroll, col_bytes=col_bytes, use_graphs=graphs, wide=wide)
            x, y = x0.copy(), y0.copy()
            r = ctx.solve_both(x, y, 1e-9, 100000, precision=prec, stop_rule=KT_STOP_TRUE)
            assert max(r["true_resid"]) <= 1e-9
            out.append((x, y))
        for x, y in out[1:]:
            assert np.array_equal(x, out[0][0]) and np.array_equal(y, out[0][1])
    ctx.close()


def test_launch_shape_setters_reject_bad_values(gpu_ctx_factory):
    """The entry-block shape setters accept only the shapes that are built and
    fail with a message otherwise; a rejected call leaves the context usable."""
    from ktplace_amd import simpl
    from ktplace_amd.simpl import KT_SPMV_ENTRY_BLOCKS, KtError, lib
    ctx = gpu_ctx_factory()
    bad = (dict(block_threads=128), dict(block_threads=256, block_entries=8192),
           dict(block_threads=512, block_entries=2048), dict(block_threads=512, unroll=8),
           dict(block_threads=1024, block_entries=2048), dict(row_lanes=2), dict(unroll=2),
           dict(block_entries=1000), dict(col_bytes=2), dict(col_bytes=5))
    for kw in bad:
        with pytest.raises(KtError):
            ctx.set_tuning(lanes=4, tile=2048, layout=KT_SPMV_ENTRY_BLOCKS, **kw)
    for kw in (dict(block_threads=256, block_entries=2048, col_bytes=3), dict(block_threads=512, block_entries=4096),
               dict(block_threads=1024, block_entries=16384, unroll=8)):
        ctx.set_tuning(lanes=4, tile=2048, layout=KT_SPMV_ENTRY_BLOCKS, **kw)
    rng = np.random.default_rng(11)
    s = spd_system(rng)
    ctx.set_system(0, s)
    ctx.set_system(1, s)
    x, y = np.zeros(s.n), np.zeros(s.n)
    r = ctx.solve_both(x, y, 1e-10, 100000)
    assert max(r["resid"]) <= 1e-10
    b = simpl.Batch(0)
    for v in (2, 5):
        assert lib.kt_batch_set_columns(b.h, v) != 0
    assert lib.kt_batch_set_columns(b.h, 3) == 0
    b.close()
    ctx.close()
