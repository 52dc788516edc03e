"""Learned warm start of the x/y solve on the GPU.

Checks: the features match a numpy restatement from the oracle's systems and
occupancy; the line search never raises the quadratic energy of the start;
an all-zero network leaves the solve exactly as from the previous iterate;
the trained network (ktplace_amd/warm_start_mlp.npy) cuts CG iterations on a
netlist it was not trained on, and both starts reach the same tolerance.
"""
import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz
from ktplace_amd import simpl

pytestmark = pytest.mark.gpu


def start_point(d, ctx, o=None, tol=1e-6):
    """seedUniform + one alpha = 0 round, then the systems of round two."""
    x, y = dz.seed_uniform(d, 20240607)
    ctx.build_b2b(x, y)
    ctx.solve_both(x, y, tol, 1_000_000)
    ctx.build_b2b(x, y)
    if o is not None:
        o.build_b2b(x, y, floors=d.flat.floors)
    return x, y


def energy(s, x):
    A = s.scipy()
    return 0.5 * x @ (A @ x) - s.rhs @ x


def numpy_features(d, o, x, y, alpha=0.0, ax=None, ay=None):
    n = d.n
    sx_, sy_ = o.system(0), o.system(1)
    rx = sx_.rhs - o.multiply(0, x)
    ry = sy_.rhs - o.multiply(1, y)
    cx, cy = x.mean(), y.mean()
    W, H = d.die[2] - d.die[0], d.die[3] - d.die[1]
    s_x, s_y = max(x.std(), 1e-6 * W), max(y.std(), 1e-6 * H)
    f = np.zeros((n, 17))
    f[:, 0] = (x - cx) / s_x
    f[:, 1] = (y - cy) / s_y
    f[:, 2] = 2 * (x - (d.die[0] + d.die[2]) / 2) / W
    f[:, 3] = 2 * (y - (d.die[1] + d.die[3]) / 2) / H
    f[:, 4] = s_x / W
    f[:, 5] = s_y / H
    ps = d.flat.pin_start.astype(np.int64)
    f[:, 6] = np.log2(1 + np.diff(ps)) / 4
    f[:, 7] = np.clip(np.log(sx_.diag / sx_.diag.mean()) / 4, -2, 2)
    f[:, 8] = np.clip(np.log(sy_.diag / sy_.diag.mean()) / 4, -2, 2)
    occ = o.occupancy(x, y)
    avail, _ = o.grid_avail()
    nb = o.bins
    gx = np.clip((x + 0.5 * d.flat.width - d.die[0]) / ((d.die[2] - d.die[0]) / nb), 0, nb - 1e-9).astype(int)
    gy = np.clip((y + 0.5 * d.flat.height - d.die[1]) / ((d.die[3] - d.die[1]) / nb), 0, nb - 1e-9).astype(int)
    b = gy * nb + gx
    with np.errstate(divide="ignore", invalid="ignore"):
        ov = np.where(avail[b] > 0, np.clip(occ[b] / avail[b] / 4, 0, 2), np.where(occ[b] > 0, 2.0, 0.0))
    f[:, 9] = ov
    f[:, 10] = np.clip(rx / sx_.diag / s_x, -4, 4)
    f[:, 11] = np.clip(ry / sy_.diag / s_y, -4, 4)
    if alpha > 0:
        f[:, 12] = np.clip((ax - x) / s_x, -4, 4)
        f[:, 13] = np.clip((ay - y) / s_y, -4, 4)
        f[:, 14] = min(alpha, 4) / 4
    for k, s in ((15, sx_), (16, sy_)):
        off = np.zeros(n)
        np.add.at(off, np.repeat(np.arange(n), np.diff(s.row_ptr.astype(np.int64))), np.abs(s.val))
        f[:, k] = np.clip((s.diag - off) / s.diag, 0, 1)
    return f, (s_x, s_y)


@pytest.mark.parametrize("anchored", [False, True])
def test_features_match_numpy(gpu_ctx_factory, anchored):
    d = dz.workload("cells12k")
    o = ol.Oracle.of(d)
    ctx = gpu_ctx_factory(d)
    ctx.build_grid(0, 1.0)
    o.build_grid(0, 1.0)
    x, y = start_point(d, ctx, o)
    alpha, ax, ay = 0.0, None, None
    if anchored:
        ax, ay = ctx.lal(x, y)
        alpha = 0.3
        ctx.build_b2b(x, y, alpha, ax, ay)
        o.build_b2b(x, y, alpha, ax, ay, floors=d.flat.floors)
    ctx.set_positions(x, y)
    g, (gs_x, gs_y) = ctx.warm_features()
    ref, (s_x, s_y) = numpy_features(d, o, x, y, alpha, ax, ay)
    assert gs_x == pytest.approx(s_x, rel=1e-9) and gs_y == pytest.approx(s_y, rel=1e-9)
    for k in range(17):
        assert np.allclose(g[:, k], ref[:, k], rtol=2e-6, atol=2e-6), k
    ctx.close()


def test_zero_network_is_the_previous_iterate_start(gpu_ctx_factory):
    d = dz.workload("cells12k")
    ctx = gpu_ctx_factory(d)
    ctx.build_grid(0, 1.0)
    x, y = start_point(d, ctx)
    ctx.set_warm_start(np.zeros(simpl.KT_MLP_PARAMS, np.float32))
    a = x.copy(), y.copy()
    ra = ctx.solve_both(*a, 1e-6, 100000, stop_rule=simpl.KT_STOP_TRUE)
    ctx.build_b2b(x, y)
    b = x.copy(), y.copy()
    rb = ctx.solve_both(*b, 1e-6, 100000, stop_rule=simpl.KT_STOP_TRUE, warm_start=simpl.KT_WARM_LEARNED)
    assert rb["warm_gamma"] == (0.0, 0.0)
    assert ra["iterations"] == rb["iterations"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ctx.close()


def test_line_search_never_raises_energy(gpu_ctx_factory):
    """Whatever the network predicts, the start's energy x'Ax/2 - b'x is at
    most the previous iterate's, and the solve still reaches the tolerance."""
    from ktplace_amd.learned import init_params
    d = dz.workload("cells12k")
    o = ol.Oracle.of(d)
    ctx = gpu_ctx_factory(d)
    ctx.build_grid(0, 1.0)
    x, y = start_point(d, ctx, o)
    sx, sy = o.system(0), o.system(1)
    e0 = energy(sx, x), energy(sy, y)
    for seed in (1, 2, 3):
        ctx.set_warm_start(init_params(seed, out_scale=3.0))
        ctx.build_b2b(x, y)  # positions and systems of the start point
        dx, dy = ctx.warm_predict()
        assert np.all(np.isfinite(dx)) and np.abs(dx).max() > 0
        g = []
        for s, v, dv, e in ((sx, x, dx, e0[0]), (sy, y, dy, e0[1])):
            A = s.scipy()
            r = s.rhs - A @ v
            gam = max((dv @ r) / (dv @ (A @ dv)), 0.0)
            assert energy(s, v + gam * dv) <= e + 1e-9 * abs(e)
            g.append(gam)
        xs, ys = x.copy(), y.copy()
        ctx.build_b2b(x, y)
        r = ctx.solve_both(xs, ys, 1e-6, 100000, stop_rule=simpl.KT_STOP_TRUE, warm_start=simpl.KT_WARM_LEARNED)
        print("seed", seed, "gamma gpu", r["warm_gamma"], "numpy", g, flush=True)
        assert r["warm_gamma"][0] == pytest.approx(g[0], rel=1e-6, abs=1e-12)
        assert r["warm_gamma"][1] == pytest.approx(g[1], rel=1e-6, abs=1e-12)
        assert max(r["true_resid"]) <= 1e-6
    ctx.close()


def test_trained_network_cuts_iterations_on_unseen_netlist(gpu_ctx_factory):
    from ktplace_amd.learned import DEFAULT_PARAMS, ab_test, load_params
    if not DEFAULT_PARAMS.exists():
        pytest.skip("no trained parameters")
    d = dz.generate(150_000, 165_000, 300, seed=777)  # not among the training seeds
    ctx = gpu_ctx_factory()
    for prec in (simpl.KT_PREC_FP64, simpl.KT_PREC_MIXED):
        ab = ab_test(d, ctx, load_params(), precision=prec)
        print(prec, ab, flush=True)
        assert max(ab["previous"]["true_resid"]) <= 1e-6 and max(ab["learned"]["true_resid"]) <= 1e-6
        assert ab["learned"]["iterations"] < ab["previous"]["iterations"]
        assert min(ab["learned"]["gamma"]) > 0
    ctx.close()


def floating_cells_scipy(d):
    """Movable cells whose connected component (vertices joined by nets) holds
    no fixed vertex, from scipy's connected components of the pin graph."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    f = d.flat
    V, N = f.num_vertices, f.num_nets
    net_of_pin = np.repeat(np.arange(N), np.diff(f.net_start.astype(np.int64)))
    g = sp.csr_matrix((np.ones(len(f.pin_cell)), (f.pin_cell.astype(np.int64), net_of_pin)), shape=(V, N))
    _, lab = connected_components(sp.bmat([[None, g], [g.T, None]]).tocsr(), directed=False)
    fixed = f.var_of_vertex == 0xFFFFFFFF
    anchored = np.zeros(lab.max() + 1, bool)
    anchored[lab[:V][fixed]] = True
    return ~anchored[lab[:V][f.mov_vertex.astype(np.int64)]]


def test_learned_start_leaves_floating_components(gpu_ctx_factory):
    """Cells with no net path to a pad or macro form components whose
    translation is in the null space of A when there are no anchors: the
    learned start must not move them (no solve would bring them back, and the
    solution would then differ from the reference's there by the move)."""
    from ktplace_amd.learned import DEFAULT_PARAMS, load_params
    if not DEFAULT_PARAMS.exists():
        pytest.skip("no trained parameters")
    d = dz.generate(150_000, 165_000, 300, seed=777)
    want = floating_cells_scipy(d)
    ctx = gpu_ctx_factory(d)
    count, flags = ctx.floating_cells()
    assert count == int(want.sum()) > 0 and np.array_equal(flags.astype(bool), want)
    ctx.build_grid(0, 1.0)
    ctx.set_warm_start(load_params())
    x, y = dz.seed_uniform(d, 20240607)
    ctx.build_b2b(x, y)
    ctx.solve_both(x, y, 1e-6, 100000)
    ctx.build_b2b(x, y)
    sx, sy = x.copy(), y.copy()
    r = ctx.solve_both(sx, sy, 1e-6, 0, warm_start=simpl.KT_WARM_LEARNED)  # the start only
    assert min(r["warm_gamma"]) > 0
    assert np.array_equal(sx[want], x[want]) and np.array_equal(sy[want], y[want])
    assert np.mean(sx[~want] != x[~want]) > 0.9
    ctx.close()
