"""Batched small designs: many uneven netlists' x/y solves in one launch.

Workload: `count` seeded synthetic netlists (default 256) whose movable cell
counts are log-uniform in [2,000, 40,000], with the degree mix of the
12.5K-cell workload (nets = 1.12 x cells, pads = cells / 50, at least 40, 70 %
utilisation).  Each instance's systems are the reference's single-solve
start: seedUniform, one alpha = 0 B2B round solved to the tolerance, then the
B2B systems at that point (simpl_placer.cc:419-424).

`run` solves all instances to the true relative residual `tol` twice:
  batched  one kt_batch: every CG kernel launch covers all tiles of all
           systems; iteration pairs replay one captured hipGraph; launch shape
           (lanes per row, tile, rows in flight) from `autotune`;
  loop     one context per instance, solved one after the other with the
           two-system solver (its own graph replay, either SpMV layout, the
           faster reported), the per-instance launch loop.
and reports solves per second and the p50/p99 per-instance solve time.  In the
batched run an instance's time is when the host poll first saw both of its
systems stopped.

    python -m ktplace_amd.batched --count 256
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import design as dz
from . import simpl

LANES = (4, 8, 16)
TILES = (256, 512, 1024, 2048)
ROWS = (2, 4)


def sizes(count: int, lo: int = 2000, hi: int = 40000, seed: int = 5000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(int)


def make_designs(count: int, seed: int = 5000):
    out = []
    for k, n in enumerate(sizes(count, seed=seed)):
        n = int(n)
        out.append(dz.generate(n, int(round(n * 1.12)), max(40, n // 50), seed=seed + 1 + k))
    return out


def start_context(d, tol: float) -> simpl.Context:
    """A context holding `d`'s start-point systems and positions (saved)."""
    ctx = simpl.Context(0).upload(d)
    x, y = dz.seed_uniform(d, 20240607)
    ctx.build_b2b(x, y)
    ctx.solve_both(x, y, tol, 1_000_000)
    ctx.build_b2b(x, y)
    ctx.save_positions()
    return ctx


def autotune(batch: simpl.Batch, iterations: int = 48):
    """Fastest SpMV layout and (lanes, tile, rows) for `iterations` lock-step
    iterations of the whole batch."""
    table = []

    def timed(g, t, r, layout=0, row_lanes=0, col_bytes=0, wide=0):
        batch.set_tuning(lanes=g, tile=t, rows=r, layout=layout, row_lanes=row_lanes, col_bytes=col_bytes,
                         wide=wide)
        batch.reset()
        batch.solve(0.0, iterations)  # warm-up (captures the graph)
        batch.reset()
        t0 = time.perf_counter()
        batch.solve(0.0, iterations)
        table.append(((time.perf_counter() - t0) * 1e3, g, t, r, layout, row_lanes, col_bytes, wide))

    for g in LANES:
        for t in TILES:
            for r in ROWS:
                timed(g, t, r)
    for t in TILES:  # entry blocks: lanes and rows do not apply; the tile shapes the vector kernels
        for row_lanes in (1, 4):
            for col_bytes in (4, 3):
                for wide in ((0, 1) if row_lanes == 4 else (0,)):  # wide loads need 4 row lanes
                    timed(4, t, 2, simpl.KT_SPMV_ENTRY_BLOCKS, row_lanes, col_bytes, wide)
    table.sort()
    ms, g, t, r, layout, row_lanes, col_bytes, wide = table[0]
    batch.set_tuning(lanes=g, tile=t, rows=r, layout=layout, row_lanes=row_lanes, col_bytes=col_bytes, wide=wide)
    return dict(lanes_per_row=g, tile_rows=t, rows_in_flight=r,
                spmv_layout="entry blocks" if layout else "row groups", row_lanes=row_lanes,
                column_bytes=(col_bytes if layout else None), wide_loads=(bool(wide) if layout else None),
                ms_per_48_iterations=round(ms, 3))


def percentiles(v):
    v = np.asarray(v, np.float64)
    return float(np.percentile(v, 50)), float(np.percentile(v, 99))


def run(count: int = 256, steps: int = 3, tol: float = 1e-6, seed: int = 5000, log=print) -> dict:
    t0 = time.perf_counter()
    designs = make_designs(count, seed)
    log(f"[batched] generated {count} netlists ({time.perf_counter() - t0:.1f} s)")
    ctxs = [start_context(d, tol) for d in designs]
    batch = simpl.Batch(0)
    for c in ctxs:
        batch.add_ctx(c)
    inst, rows, nnz = batch.size()
    log(f"[batched] {inst} instances, {rows} rows, {nnz} entries ({time.perf_counter() - t0:.1f} s)")
    tuning = autotune(batch)
    log(f"[batched] tuning {tuning}")

    # batched: one kt_batch_solve per step
    batch.reset()
    batch.solve(tol, 1_000_000, simpl.KT_STOP_TRUE)
    walls, per = [], None
    for _ in range(steps):
        batch.reset()
        t1 = time.perf_counter()
        per, tot = batch.solve(tol, 1_000_000, simpl.KT_STOP_TRUE)
        walls.append((time.perf_counter() - t1) * 1e3)
    b_ms = float(np.median(walls))
    b50, b99 = percentiles([p["done_ms"] for p in per])
    worst = max(max(p["true_resid"]) for p in per)
    log(f"[batched] {b_ms:.2f} ms per batch, p50 {b50:.2f} p99 {b99:.2f} ms, worst true resid {worst:.2e}")

    # loop: each instance alone, one after the other, with the two-system
    # solver in either SpMV layout, entry blocks of 1024 threads or of 256
    # threads with 3-byte columns (the fastest is reported)
    best = None
    for layout, threads, col_bytes in ((simpl.KT_SPMV_ROW_GROUPS, 0, 0), (simpl.KT_SPMV_ENTRY_BLOCKS, 0, 0),
                                       (simpl.KT_SPMV_ENTRY_BLOCKS, 256, 3)):
        name = ("row groups" if not layout else
                f"entry blocks of {threads or 1024} threads, {col_bytes or 4}-byte columns")
        for c in ctxs:  # warm-up (graph capture per context)
            c.set_tuning(layout=layout, block_threads=threads, col_bytes=col_bytes)
            c.restore_positions()
            c.solve_resident(tol, 1_000_000, stop_rule=simpl.KT_STOP_TRUE)
        loop_totals, loop_each = [], []
        for _ in range(steps):
            each = []
            for c in ctxs:
                c.restore_positions()
                c.sync()
                t1 = time.perf_counter()
                c.solve_resident(tol, 1_000_000, stop_rule=simpl.KT_STOP_TRUE)
                each.append((time.perf_counter() - t1) * 1e3)
            loop_totals.append(sum(each))
            loop_each = each
        ms = float(np.median(loop_totals))
        log(f"[batched] loop ({name}) {ms:.2f} ms per pass")
        if best is None or ms < best[0]:
            best = (ms, loop_each, name)
    l_ms, loop_each, loop_layout = best
    l50, l99 = percentiles(loop_each)
    log(f"[batched] loop {l_ms:.2f} ms per pass, p50 {l50:.2f} p99 {l99:.2f} ms")
    iters = [max(p["steps"]) for p in per]
    out = dict(
        instances=inst, cells_min=int(min(d.n for d in designs)), cells_max=int(max(d.n for d in designs)),
        rows=rows, nnz=nnz, tol=tol, stop="true FP64 relative residual",
        iterations_p50=float(np.median(iters)), iterations_max=int(max(iters)),
        batched=dict(ms=round(b_ms, 3), solves_per_s=round(inst / (b_ms * 1e-3), 1), p50_ms=round(b50, 3),
                     p99_ms=round(b99, 3), worst_true_resid=worst, tuning=tuning),
        loop=dict(ms=round(l_ms, 3), solves_per_s=round(inst / (l_ms * 1e-3), 1), p50_ms=round(l50, 3),
                  p99_ms=round(l99, 3), spmv_layout=loop_layout),
        speedup=round(l_ms / b_ms, 2))
    batch.close()
    for c in ctxs:
        c.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=5000)
    args = ap.parse_args(argv)
    res = run(args.count, args.steps, args.tol, args.seed,
              log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
