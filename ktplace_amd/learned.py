"""Learned warm start: data from SimPL runs, training on the GPU, parameters.

The network (libktplace_amd: csrc/mlp.hip) predicts each cell's displacement
from its previous iterate to the solution of the newly built B2B systems, in
units of the placement's spread; the solve applies it through an exact line
search that falls back to the previous iterate (csrc/warm.hip).  The
reference has no such part; its solves start from the previous iterate
(simpl_lss.cc:214-222).

Training data are (features, displacement) pairs that SimPL itself generates:
`collect` runs the placement loop of placer.Placer with every solve taken to a
tight tolerance, and records, at every solve, the features at the previous
iterate and the displacement to the solution, for a sample of cells.  `main`
generates seeded netlists, collects, trains with the GPU trainer and writes
the parameters next to this file (warm_start_mlp.npy, with a JSON summary).

    python -m ktplace_amd.learned --out ktplace_amd/warm_start_mlp.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib
from . import simpl
from .simpl import KT_MLP_FEATURES, KT_MLP_PARAMS, _check, lib

HERE = Path(__file__).resolve().parent
DEFAULT_PARAMS = HERE / "warm_start_mlp.npy"
FEATURE_NAMES = ("x_rel", "y_rel", "x_die", "y_die", "spread_x", "spread_y", "pins", "log_diag_x",
                 "log_diag_y", "bin_density", "jacobi_x", "jacobi_y", "anchor_x", "anchor_y", "alpha",
                 "fixed_share_x", "fixed_share_y")
SHAPES = (("W1", (128, 32)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
          ("W3", (2, 128)), ("b3", (2,)))


def init_params(seed: int, out_scale: float = 0.1) -> np.ndarray:
    """He-uniform weights and zero biases; the output layer is scaled down so
    that the untrained network predicts displacements near zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = math.sqrt(6.0 / shape[1])
            w = rng.uniform(-lim, lim, shape)
            if name == "W1":
                w[:, KT_MLP_FEATURES:] = 0.0  # inputs that are always zero
            parts.append(w.reshape(-1) * (out_scale if name == "W3" else 1.0))
        else:
            parts.append(np.zeros(shape))
    p = np.concatenate(parts).astype(np.float32)
    assert len(p) == KT_MLP_PARAMS
    return p


def load_params(path: Optional[Path] = None) -> np.ndarray:
    path = Path(path) if path else DEFAULT_PARAMS
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (KT_MLP_PARAMS,):
        raise ValueError(f"{path}: expected {KT_MLP_PARAMS} float32 parameters")
    return p


class Mlp:
    """The GPU trainer (kt_mlp): forward, gradient and Adam steps."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib.kt_mlp_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.kt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib.kt_mlp_set_params(self.h, _lib.ptr(p), len(p)))

    def get_params(self) -> np.ndarray:
        p = np.zeros(KT_MLP_PARAMS, np.float32)
        _check(lib.kt_mlp_get_params(self.h, _lib.ptr(p), len(p)))
        return p

    def set_data(self, features, targets, weights=None):
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        if f.shape[1] != KT_MLP_FEATURES or t.shape != (len(f), 2) or (w is not None and w.shape != (len(f),)):
            raise ValueError("features [n][17], targets [n][2], weights [n]")
        _check(lib.kt_mlp_set_data(self.h, _lib.ptr(f), _lib.ptr(t), _lib.ptr(w), len(f)))
        self._data = (f, t, w)

    def forward(self, features) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32)
        out = np.zeros((len(f), 2), np.float32)
        _check(lib.kt_mlp_forward(self.h, _lib.ptr(f), len(f), _lib.ptr(out)))
        return out

    def gradient(self, samples):
        s = np.ascontiguousarray(samples, np.uint32)
        g = np.zeros(KT_MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib.kt_mlp_gradient(self.h, _lib.ptr(s), len(s), _lib.ptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, steps: int, batch: int, lr: float, seed: int) -> np.ndarray:
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib.kt_mlp_train(self.h, steps, batch, lr, seed, _lib.ptr(loss)))
        return loss[:steps]


# --------------------------------------------------------------- data


def collect(design, ctx: simpl.Context, cells_per_solve: int, rng: np.random.Generator,
            max_global: int = 12, tol: float = 1e-6):
    """Runs SimPL on `design` with tight solves and returns the pairs it
    generates: (features [m][17], targets [m][2], solve id [m], solve kinds)."""
    from .placer import Placer, SimplOptions

    feats, targs, ids, kinds = [], [], [], []

    class Recorder(Placer):
        def _solve(self, x, y):
            self.ctx.set_positions(x, y)
            f, (sx, sy) = self.ctx.warm_features()
            x0, y0 = x.copy(), y.copy()
            r = super()._solve(x, y)
            t = np.stack([(x - x0) / sx, (y - y0) / sy], 1)
            k = min(cells_per_solve, len(x))
            pick = rng.choice(len(x), k, replace=False)
            feats.append(f[pick])
            targs.append(t[pick].astype(np.float32))
            ids.append(np.full(k, len(kinds), np.int32))
            kinds.append("anchored" if f[0, 14] > 0 else "initial")
            return r

    opts = SimplOptions(initialNetModel="b2b", cgTol=tol, cgMaxIter=100_000, maxIters=max_global,
                        refineIters=0, precision=simpl.KT_PREC_FP64)
    rec = Recorder(design, opts, ctx)
    rec.run()
    return np.concatenate(feats), np.concatenate(targs), np.concatenate(ids), kinds


def solve_weights(targets: np.ndarray, ids: np.ndarray) -> np.ndarray:
    """Per-sample loss weights that give every solve the same total weight
    relative to its displacement: 1 / mean |t|^2 over the solve's samples."""
    w = np.zeros(len(ids), np.float32)
    for s in np.unique(ids):
        m = ids == s
        e = float(np.mean(np.sum(targets[m].astype(np.float64) ** 2, 1)))
        w[m] = 1.0 / max(e, 1e-12)
    return w / w.mean()


def ab_test(design, ctx: simpl.Context, params: np.ndarray, tol: float = 1e-6,
            precision: int = simpl.KT_PREC_FP64) -> dict:
    """CG iterations to `tol` (true residual) on the second initial B2B round
    of `design` (start: seedUniform + one alpha = 0 round) from the previous
    iterate and from the learned start."""
    from .design import seed_uniform
    ctx.upload(design)
    ctx.build_grid(0, 1.0)
    x0, y0 = seed_uniform(design, 20240607)
    ctx.build_b2b(x0, y0)
    ctx.solve_both(x0, y0, tol, 1_000_000)
    out = {}
    ctx.set_warm_start(params)
    for name, warm in (("previous", simpl.KT_WARM_PREVIOUS), ("learned", simpl.KT_WARM_LEARNED)):
        x, y = x0.copy(), y0.copy()
        ctx.build_b2b(x, y)
        r = ctx.solve_both(x, y, tol, 1_000_000, precision=precision, stop_rule=simpl.KT_STOP_TRUE,
                           warm_start=warm)
        out[name] = dict(iterations=r["iterations"], steps=r["steps"], true_resid=r["true_resid"],
                         gamma=r["warm_gamma"])
    return out


def main(argv=None) -> int:
    from . import design as dz
    ap = argparse.ArgumentParser(description="train the warm-start network on SimPL-generated pairs")
    ap.add_argument("--netlists", type=int, default=64)
    ap.add_argument("--min-cells", type=int, default=50_000)
    ap.add_argument("--max-cells", type=int, default=200_000)
    ap.add_argument("--holdout", type=int, default=4, help="netlists kept out of training")
    ap.add_argument("--seed", type=int, default=1000, help="first netlist seed")
    ap.add_argument("--cells-per-solve", type=int, default=4096)
    ap.add_argument("--max-global", type=int, default=12)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=str(DEFAULT_PARAMS))
    args = ap.parse_args(argv)

    rng = np.random.default_rng(args.seed)
    sizes = np.exp(rng.uniform(math.log(args.min_cells), math.log(args.max_cells), args.netlists)).astype(int)
    ctx = simpl.Context(0)
    F, T, W, meta = [], [], [], []
    val = []
    t0 = time.perf_counter()
    for k, n in enumerate(sizes):
        seed = args.seed + k
        d = dz.generate(int(n), int(round(n * 1.1)), max(250, int(n) // 500), seed=seed)
        f, t, ids, kinds = collect(d, ctx, args.cells_per_solve, rng, args.max_global)
        w = solve_weights(t, ids)
        part = "holdout" if k >= args.netlists - args.holdout else "train"
        (val if part == "holdout" else F).append((f, t, w, d) if part == "holdout" else f)
        if part == "train":
            T.append(t)
            W.append(w)
        meta.append(dict(seed=seed, cells=int(n), solves=len(kinds), part=part))
        print(f"[learned] netlist {k} seed {seed} cells {n}: {len(kinds)} solves, {len(f)} samples "
              f"({part}, {time.perf_counter() - t0:.0f} s)", flush=True)
    feats = np.concatenate(F)
    targs = np.concatenate(T)
    weights = np.concatenate(W)
    m = Mlp(0)
    p = init_params(args.seed)
    m.set_params(p)
    m.set_data(feats, targs, weights)
    losses = []
    # constant learning rate for half the steps, then linear decay to 5 %
    chunk = 1000
    for s0 in range(0, args.steps, chunk):
        frac = s0 / args.steps
        lr = args.lr if frac < 0.5 else args.lr * max(0.05, 2.0 * (1.0 - frac))
        losses.extend(m.train(min(chunk, args.steps - s0), args.batch, lr, args.seed + s0).tolist())
        print(f"[learned] step {s0 + chunk}: lr {lr:.2e} loss {np.mean(losses[-chunk:]):.4f}", flush=True)
    params = m.get_params()
    # held-out loss and the A/B test on held-out netlists
    holdout = []
    for f, t, w, d in val:
        out = m.forward(f)
        vloss = float(np.mean(w * np.sum((out - t) ** 2, 1)))
        base = float(np.mean(w * np.sum(t.astype(np.float64) ** 2, 1)))
        ab = ab_test(d, ctx, params)
        holdout.append(dict(cells=d.n, loss=vloss, zero_prediction_loss=base, ab=ab))
        print(f"[learned] holdout {d.n} cells: loss {vloss:.4f} (zero prediction {base:.4f}); "
              f"iterations previous {ab['previous']['iterations']} learned {ab['learned']['iterations']} "
              f"gamma {ab['learned']['gamma']}", flush=True)
    out = Path(args.out)
    np.save(out, params, allow_pickle=False)
    summary = dict(netlists=meta, samples=int(len(feats)), steps=args.steps, batch=args.batch, lr=args.lr,
                   features=FEATURE_NAMES, final_loss=float(np.mean(losses[-200:])), holdout=holdout,
                   seconds=time.perf_counter() - t0)
    out.with_suffix(".json").write_text(json.dumps(summary, indent=1))
    print(f"[learned] wrote {out} ({time.perf_counter() - t0:.0f} s)", flush=True)
    m.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
