#!/usr/bin/env python3
"""B2B + Jacobi-PCG solve time on one MI355X, against the CPU oracle.

One step = one B2B assembly of the x and y systems plus their lock-step
Jacobi-PCG solve until the true FP64 relative residual ||b - Ax|| / ||b|| is
at most 1e-6 on both axes, on a seeded synthetic netlist resident in HBM:
1,000,000 movable cells, 1,100,000 nets (degree 2: 60 %, 3: 20 %, 4-10: 15 %,
11-40: 4.9 %, 41-200: 0.1 %), 2,000 perimeter pads.  The start point is the
reference's: seedUniform followed by one alpha = 0 B2B round (untimed setup);
every step restores it before building, so each step does the same work.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "B2B+PCG solve ms to 1e-6 rel residual @1M cells: mixed+learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
# HBM bytes per SpMV dispatch on this workload (averaged over the dispatches of a
# solve), from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by the copy-probe
# calibration: profiles/cgSpmv_f64_cells1m.md, profiles/cgSpmv_f32_cells1m.md
MEASURED_TRAFFIC = {"fp64": int((552.0 + 13.3) * 1e6), "mixed": int((298.2 + 7.6) * 1e6)}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def spmv_bytes(nnz: int, n: int, precision: str, fused: bool = False) -> int:
    """Algorithmic bytes of one SpMV over one axis: col (4) + val per entry;
    rowPtr (4) + diag + x + y per row.  The fused FP32 iteration's SpMV also
    reads the residual r (4 bytes per row) for its r.Ap and r.r partials."""
    if precision == "fp64":
        return 12 * nnz + 28 * n
    return 8 * nnz + 16 * n + (4 * n if fused else 0)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(d, x1, y1, tol: float, runs: int = 5) -> dict:
    """The oracle (CPU restatement of the reference) on the same start point,
    timed on this host's cores.

    All cores: one warm-up step, then the median of `runs` full steps (B2B
    build of both axes + FP64 Jacobi-PCG to tol, the reference's recurrence
    stopping rule, iteration cap lifted).  One core: a bounded sample, the
    median of 3 runs of the build plus 60 CG iterations (the reference's
    default cap), scaled to the full step's iteration count."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol
    nproc = os.cpu_count() or 1
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc))
    threads = max(1, min(threads, nproc))
    o = ol.Oracle.of(d)

    def full_step():
        t0 = time.perf_counter()
        o.build_b2b(x1, y1, floors=d.flat.floors)
        t1 = time.perf_counter()
        r = o.solve(x1, y1, tol=tol, max_iter=1_000_000)
        return 1e3 * (t1 - t0), 1e3 * (time.perf_counter() - t1), r["iterations"]

    ol.set_threads(threads)
    full_step()  # warm-up
    all_runs = [full_step() for _ in range(runs)]
    builds, solves, iters = zip(*all_runs)
    total = [b + s for b, s, _ in all_runs]
    iterations = iters[-1]

    ol.set_threads(1)
    sample = 60
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.build_b2b(x1, y1, floors=d.flat.floors)
        t1 = time.perf_counter()
        o.solve(x1, y1, tol=0.0, max_iter=sample)
        t2 = time.perf_counter()
        one.append((1e3 * (t1 - t0), 1e3 * (t2 - t1) / sample))
    one_build = float(np.median([b for b, _ in one]))
    one_iter = float(np.median([i for _, i in one]))
    ol.set_threads(threads)
    return dict(
        value=float(np.median(total)), unit="ms", cores=threads, kind="port",
        sample=(f"the oracle on the same {d.n}-cell system and start point: one B2B build of both axes + "
                f"FP64 Jacobi-PCG to {tol:g} (recurrence rule), {iterations} iterations; median of {runs} "
                f"runs after one warm-up on {threads} threads"),
        build_ms=round(float(np.median(builds)), 1), solve_ms=round(float(np.median(solves)), 1),
        runs_ms=[round(v, 1) for v in total], iterations=iterations,
        one_core=dict(value=round(one_build + one_iter * iterations, 1), unit="ms", build_ms=round(one_build, 1),
                      ms_per_iteration=round(one_iter, 3),
                      sample=(f"median of 3 runs of the build + {sample} CG iterations on 1 thread, "
                              f"scaled to {iterations} iterations")),
        cpu_model=cpu_model(), nproc=nproc)


def oracle_solution(d, x1, y1) -> dict:
    """The oracle's FP64 solution of the start-point systems to a relative
    residual of 1e-12: the reference every GPU path's error is taken against
    (computed outside every timed region)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol
    o = ol.Oracle.of(d)
    o.build_b2b(x1, y1, floors=d.flat.floors)
    return o.solve(x1, y1, tol=1e-12, max_iter=1_000_000)


def rel_error(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cells1m")
    ap.add_argument("--paths", default="fp64,mixed,mixed+learned",
                    help="comma list of solver paths to time: fp64 (plain), mixed, mixed+learned, "
                         "fp64+learned")
    ap.add_argument("--batched-count", type=int, default=256,
                    help="also time the batched small-designs workload with this many netlists (0: skip)")
    ap.add_argument("--warm-params", default="",
                    help="warm-start network parameters (.npy; default: the package's trained ones)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-error", action="store_true",
                    help="skip the FP64 oracle solution that every path's error is taken against")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run the plain copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--lanes", type=int, default=0, help="skip the autotuner: lanes per row")
    ap.add_argument("--tile", type=int, default=0, help="skip the autotuner: rows per block")
    ap.add_argument("--xcd-split", type=int, default=1, help="with --lanes/--tile: 1 axis per XCD half")
    ap.add_argument("--rows", type=int, default=2, help="with --lanes/--tile: rows in flight per lane group")
    ap.add_argument("--halo", type=int, default=0, help="with --lanes/--tile: SpMV LDS window halo (entries)")
    ap.add_argument("--layout", type=int, default=0, help="with --lanes/--tile: 1 = entry-block SpMV")
    ap.add_argument("--block-entries", type=int, default=0, help="with --layout 1: entries per block")
    ap.add_argument("--unroll", type=int, default=0, help="with --layout 1: entries per thread and round")
    ap.add_argument("--row-lanes", type=int, default=0, help="with --layout 1: lanes per row in the row sums")
    ap.add_argument("--block-threads", type=int, default=0, help="with --layout 1: threads per block")
    ap.add_argument("--col-bytes", type=int, default=0, help="with --layout 1: bytes per column index (3 or 4)")
    ap.add_argument("--wide", type=int, default=0, help="with --layout 1: 1 = four consecutive entries per load")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the x/y systems and batched designs would shard with no exchange; see the "
              "design notes)", file=sys.stderr)
        return 2

    from ktplace_amd import design as dz
    from ktplace_amd import simpl

    t = time.perf_counter()
    d = dz.workload(args.workload, seed=args.seed)
    log(f"[bench] generated {args.workload}: n={d.n} nets={d.flat.num_nets} pins={d.flat.num_pins} "
        f"({time.perf_counter() - t:.1f} s)")
    ctx = simpl.Context(0).upload(d)

    # start point: seedUniform + one alpha = 0 B2B round (untimed)
    x0, y0 = dz.seed_uniform(d, 20240607)
    ctx.build_b2b(x0, y0)
    x1, y1 = x0.copy(), y0.copy()
    r0 = ctx.solve_both(x1, y1, args.tol, 1_000_000)
    log(f"[bench] start point: round-0 solve {r0['iterations']} iterations")
    ctx.set_positions(x1, y1)
    ctx.save_positions()
    ctx.build_b2b_resident(0.0)
    nnz = [ctx.system_size(a)[1] for a in (0, 1)]
    bins = ctx.build_grid(0, 1.0)  # density grid for the warm start's bin feature
    paths = [p for p in args.paths.split(",") if p]
    if any(p.endswith("+learned") for p in paths):
        from ktplace_amd.learned import load_params
        ctx.set_warm_start(load_params(args.warm_params or None))
        log(f"[bench] warm-start network loaded; {bins}x{bins} bins")

    ref = None
    if not args.no_oracle_error:
        t = time.perf_counter()
        ref = oracle_solution(d, x1, y1)
        log(f"[bench] FP64 oracle solution to 1e-12: {ref['iterations']} iterations "
            f"({time.perf_counter() - t:.1f} s)")

    from ktplace_amd.autotune import TuneResult, autotune
    per = {}
    failed = []
    tuned_by_prec = {}
    for path in paths:
        base_path = path.split("+")[0]
        prec = simpl.KT_PREC_FP64 if base_path == "fp64" else simpl.KT_PREC_MIXED
        warm = simpl.KT_WARM_LEARNED if path.endswith("+learned") else simpl.KT_WARM_PREVIOUS
        ctx.restore_positions()
        ctx.build_b2b_resident(0.0)
        if args.lanes and args.tile:
            tuned = TuneResult(args.lanes, args.tile, float("nan"), [], args.xcd_split, args.rows, args.halo,
                               args.layout, args.block_entries, args.unroll, args.row_lanes, args.block_threads,
                               args.col_bytes, args.wide)
        elif prec in tuned_by_prec:
            tuned = tuned_by_prec[prec]
        else:
            tuned = autotune(ctx, prec)
            tuned_by_prec[prec] = tuned
        ctx.set_tuning(**tuned.kwargs())
        log(f"[bench] {path}: autotune {tuned.kwargs()} ({tuned.ms_per_iteration:.3f} ms/iteration)")
        for cand, ms in sorted(tuned.table, key=lambda t: t[1])[:6]:
            log(f"[bench]   candidate {cand}: {ms:.4f} ms/iteration")

        def step(profile: bool, prec=prec, warm=warm) -> dict:
            ctx.restore_positions()
            ctx.build_b2b_resident(0.0)
            return ctx.solve_resident(args.tol, 1_000_000, precision=prec,
                                      stop_rule=simpl.KT_STOP_TRUE, profile=profile, warm_start=warm)

        for w in range(args.warmup):
            r = step(False)
            log(f"[bench] {path} warmup {w}: {r['iterations']} iterations, true resid {r['true_resid']}")
        # timed region: graph-replayed iterations
        ctx.sync()
        t0 = time.perf_counter()
        results = [step(False) for _ in range(args.steps)]
        ctx.sync()
        ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
        # the same steps launched one kernel at a time with HIP events around
        # each kernel (per-kernel durations for the roofline)
        ctx.reset_kernel_times()
        ctx.sync()
        t0 = time.perf_counter()
        presults = [step(True) for _ in range(args.steps)]
        ctx.sync()
        ms_prof = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
        kid = simpl.KERNEL_SPMV if base_path == "fp64" else simpl.KERNEL_SPMV32
        launches, spmv_ms = ctx.kernel_time(kid)
        # the fused FP32 iteration launches no separate p update
        fused = base_path != "fp64" and ctx.kernel_time(simpl.KERNEL_PUPDATE)[0] == 0
        total_bytes = sum(r["steps"][a] * spmv_bytes(nnz[a], d.n, base_path, fused)
                          for r in presults for a in (0, 1))
        avg_ms = spmv_ms / max(launches, 1)
        warm_n, warm_ms = ctx.kernel_time(simpl.KERNEL_WARM)
        mlp_n, mlp_ms = ctx.kernel_time(simpl.KERNEL_MLP)
        bytes_per_launch = total_bytes / max(launches, 1)
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        last = results[-1]
        for i, r in enumerate(results):
            log(f"[bench] {path} step {i}: {r['iterations']} iterations, true resid {r['true_resid']}")
        bad = [r["true_resid"] for r in results + presults if not max(r["true_resid"]) <= args.tol]
        if bad:
            failed.append(path)
            log(f"[bench] ERROR {path}: {len(bad)} timed steps ended above tol {args.tol:g}: {bad[:3]}")
        error = None
        if ref is not None:  # outside the timed region: one more step, then its positions
            step(False)
            gx, gy = ctx.get_positions()
            error = [rel_error(gx, ref["x"]), rel_error(gy, ref["y"])]
            log(f"[bench] {path}: relative error vs FP64 oracle x {error[0]:.3e} y {error[1]:.3e}")
        log(f"[bench] {path}: {ms:.2f} ms/step (graph replay), {ms_prof:.2f} ms/step with events; "
            f"SpMV {avg_ms * 1e3:.1f} us/launch, {achieved:.0f} GB/s algorithmic")
        per[path] = dict(
            ms_per_step=round(ms, 3), ms_per_step_with_kernel_events=round(ms_prof, 3),
            iterations=last["iterations"], steps_per_axis=list(last["steps"]),
            true_resid=list(last["true_resid"]),
            error_vs_fp64_oracle=error,
            refreshes_or_restarts=last["refinements"],
            warm_start=("learned (line-search gamma %.3f, %.3f)" % tuple(last["warm_gamma"])
                        if warm == simpl.KT_WARM_LEARNED else "previous iterate"),
            warm_start_ms=(round(warm_ms / warm_n, 3) if warm_n else None),
            mlp_inference_ms=(round(mlp_ms / mlp_n, 3) if mlp_n else None),
            tuning={"lanes_per_row": tuned.lanes, "tile_rows": tuned.tile,
                    "xcd_split": tuned.xcd_split, "rows_in_flight": tuned.rows, "lds_halo": tuned.halo,
                    "spmv_layout": "entry blocks" if tuned.layout else "row groups",
                    "block_threads": tuned.block_threads or (1024 if tuned.layout else None),
                    "block_entries": tuned.block_entries or (8 * (tuned.block_threads or 1024)
                                                             if tuned.layout else None),
                    "unroll": tuned.unroll or (4 if tuned.layout else None),
                    "row_lanes": tuned.row_lanes or (1 if tuned.layout else None),
                    "column_bytes": (3 if tuned.col_bytes == 3 else 4) if tuned.layout else None,
                    "wide_loads": bool(tuned.wide) if tuned.layout else None},
            roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=round(achieved / HBM_PEAK_GBS, 4),
                          traffic=(MEASURED_TRAFFIC.get(base_path) if args.workload == "cells1m" else None),
                          traffic_source="rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, cells1m, copy-probe scaled",
                          kernel=(f"{'cgSpmvFlat' if tuned.layout else 'cgSpmv'}"
                                  f"<{'double' if base_path == 'fp64' else 'float'}> (x and y per launch"
                                  f"{'; fused iteration, reads r' if fused else ''})"),
                          launches=launches, avg_launch_us=round(avg_ms * 1e3, 2),
                          bytes_per_launch=int(bytes_per_launch)))

    base = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle on the same start point ...")
        base = cpu_baseline(d, x1, y1, args.tol)
        log(f"[bench] cpu baseline {base['value']:.0f} ms on {base['cores']} threads ({base['cpu_model']}); "
            f"1 core {base['one_core']['value']:.0f} ms")

    batched = None
    if args.batched_count:
        from ktplace_amd.batched import run as run_batched
        log(f"[bench] batched small designs: {args.batched_count} netlists ...")
        batched = run_batched(args.batched_count, steps=2, tol=args.tol, log=log)

    copy = None
    if args.copy_probe:
        nb = args.copy_probe << 20
        cms = ctx.probe_copy(nb, 1)
        copy = dict(bytes=nb, ms=round(cms, 4), read_write_gbs=round(2 * nb / (cms * 1e-3) / 1e9, 1))
        log(f"[bench] copy probe {args.copy_probe} MiB: {cms:.3f} ms")

    main_path = next((p for p in ("mixed+learned", "mixed") if p in per), paths[0])
    ms = per[main_path]["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": ms,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_path.startswith("fp64") else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: one B2B assembly + lock-step Jacobi-PCG (x and y) to "
                         f"{args.tol:g} true relative residual; {d.n} cells, {d.flat.num_nets} "
                         f"nets, {d.flat.num_pins} pins, seed {args.seed}"),
            "path": main_path,
            "warm_start": per[main_path]["warm_start"],
            "paths": per,
            "speedup_vs_plain_fp64": (round(per["fp64"]["ms_per_step"] / ms, 3)
                                      if "fp64" in per else None),
            "error_vs_fp64_oracle": per[main_path]["error_vs_fp64_oracle"],
            "oracle_reference": ("the oracle's FP64 solution to a relative residual of 1e-12 "
                                 f"({ref['iterations']} iterations)" if ref is not None else None),
            "nnz_per_axis": nnz,
            "copy_probe": copy,
            "batched_designs": batched,
        },
        "roofline": per[main_path]["roofline"],
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    if failed:
        log(f"[bench] paths that missed the tolerance: {failed}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
