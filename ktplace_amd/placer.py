"""SimPL global placement driven from the host, every step on the GPU.

Mirrors simplPlacer::Impl::run (simpl_placer.cc:358-580): seedUniform, initial
rounds (star model and/or B2B, simpl_placer.cc:391-448), then outer
iterations of look-ahead legalization (upper bound), HPWL bookkeeping, the
convergence test, B2B assembly with anchors and the Jacobi-PCG solve (lower
bound), and finally the refinement on a finer grid.  Options and their
defaults follow SimplOptions (simpl_options.h:11-58).  Fences, frames and
reports are not part of this path.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import simpl
from .design import Design, seed_uniform


@dataclass
class SimplOptions:
    binsPerAxis: int = 0
    cgTol: float = 1e-4
    initMaxIters: int = 15
    initialNetModel: str = "star"  # "star", "b2b" or "star_b2b" (star rounds, then B2B rounds)
    maxIters: int = 100
    seed: int = 20240607
    densityLimit: float = 1.0
    alphaBase: float = 0.02
    maxLevel: int = 20
    minCellsToSplit: int = 4
    stripeAreaFraction: float = 0.1
    gapReferenceIter: int = 10
    gapRelaxedFrac: float = 0.25
    gapTightFrac: float = 0.05
    stallIters: int = 6
    refineIters: int = 10
    refineGridFactor: int = 3
    cgMaxIter: int = 60
    precision: int = simpl.KT_PREC_FP64


class Convergence:
    """Convergence::update (simpl_placer.cc:450-461): the gap of iteration
    gapReferenceIter is the reference; stop once the gap is below
    gapTightFrac of it, or at most gapRelaxedFrac of it while the upper bound
    has not improved for stallIters iterations."""

    def __init__(self, o: SimplOptions):
        self.o = o
        self.gap_reference = 0.0
        self.best_upper = math.inf
        self.stalled_for = 0

    def update(self, iteration: int, gap: float, upper: float) -> bool:
        o = self.o
        if iteration == o.gapReferenceIter:
            self.gap_reference = gap
        self.stalled_for = 0 if upper < self.best_upper else self.stalled_for + 1
        self.best_upper = min(self.best_upper, upper)
        stalled = self.stalled_for >= o.stallIters
        return (iteration > o.gapReferenceIter and self.gap_reference > 0.0 and
                (gap < o.gapTightFrac * self.gap_reference or
                 (gap <= o.gapRelaxedFrac * self.gap_reference and stalled)))


@dataclass
class PlaceResult:
    x: np.ndarray           # upper-bound (legalized) placement
    y: np.ndarray
    stats: dict
    trace: list = field(default_factory=list)
    lower_x: np.ndarray | None = None  # lower-bound (solved) placement
    lower_y: np.ndarray | None = None


class Placer:
    """simplPlacer over one GPU context (the design is uploaded once)."""

    def __init__(self, design: Design, opts: SimplOptions | None = None, ctx: simpl.Context | None = None,
                 progress=None):
        self.progress = progress  # called with one line per outer iteration
        self.d = design
        self.o = opts or SimplOptions()
        self.ctx = ctx or simpl.Context(0)
        self.ctx.upload(design)
        self.ctx.set_lal_options(self.o.maxLevel, self.o.minCellsToSplit, self.o.stripeAreaFraction)
        self.stats = dict(init_iters=0, global_iters=0, cg_iterations=0, build_seconds=0.0,
                          solve_seconds=0.0, spread_seconds=0.0)

    # bookkeeping (simpl_placer.cc:256-297) -----------------------------------
    def hpwl(self, x, y) -> float:
        self.ctx.set_positions(x, y)
        return self.ctx.hpwl()

    def _solve(self, x, y):
        r = self.ctx.solve_both(x, y, self.o.cgTol, self.o.cgMaxIter, precision=self.o.precision,
                                stop_rule=(simpl.KT_STOP_TRUE if self.o.precision == simpl.KT_PREC_MIXED
                                           else simpl.KT_STOP_RECURRENCE))
        self.stats["cg_iterations"] += r["iterations"]
        return r

    def _solve_resident(self):
        r = self.ctx.solve_resident(self.o.cgTol, self.o.cgMaxIter, precision=self.o.precision,
                                    stop_rule=(simpl.KT_STOP_TRUE if self.o.precision == simpl.KT_PREC_MIXED
                                               else simpl.KT_STOP_RECURRENCE))
        self.stats["cg_iterations"] += r["iterations"]
        return r

    def _legalize_into_anchors(self):
        """upper = LAL(lower) into the anchors; the positions keep the lower bound."""
        ctx = self.ctx
        ctx.save_positions()
        t0 = time.perf_counter()
        ctx.lal_resident()
        self.stats["spread_seconds"] += time.perf_counter() - t0
        ctx.anchors_from_positions()
        ctx.restore_positions()

    def _initial_rounds(self, lx, ly, hpwl_init, star: bool):
        """initialRounds (simpl_placer.cc:406-448).  Star: the star systems are
        built once with the stars at the die centre, and every round continues
        the solve from the previous round's extended solution; the first star
        round is always accepted.  B2B: the systems are rebuilt at the current
        placement every round.  A round is kept only while HPWL decreases."""
        n = self.d.n
        ex = ey = None
        if star:
            size = self.ctx.build_star()
            die = self.d.die
            ex = np.concatenate([lx, np.full(size - n, 0.5 * (die[0] + die[2]))])
            ey = np.concatenate([ly, np.full(size - n, 0.5 * (die[1] + die[3]))])
        for rnd in range(self.o.initMaxIters):
            if star:
                nx, ny = ex.copy(), ey.copy()
            else:
                nx, ny = lx.copy(), ly.copy()
                self.ctx.build_b2b(nx, ny)
            self._solve(nx, ny)
            if star:
                ex, ey = nx, ny
                nx, ny = nx[:n].copy(), ny[:n].copy()
            h = self.hpwl(nx, ny)
            self.stats["init_iters"] += 1
            first = star and rnd == 0
            if not first and not (h < hpwl_init):
                break
            lx, ly, hpwl_init = nx, ny, h
        return lx, ly, hpwl_init

    def run(self) -> PlaceResult:
        o, ctx = self.o, self.ctx
        d = self.d
        ctx.build_grid(o.binsPerAxis, o.densityLimit)
        lx, ly = seed_uniform(d, o.seed)
        self.stats["hpwl_seed"] = self.hpwl(lx, ly)
        if o.initialNetModel not in ("star", "b2b", "star_b2b"):
            raise ValueError(f"unknown initial net model {o.initialNetModel!r}")
        hpwl_init = self.stats["hpwl_seed"]
        if o.initialNetModel != "b2b":
            lx, ly, hpwl_init = self._initial_rounds(lx, ly, hpwl_init, star=True)
        if o.initialNetModel != "star":
            lx, ly, hpwl_init = self._initial_rounds(lx, ly, hpwl_init, star=False)
        self.stats["hpwl_init"] = hpwl_init
        # getGlobalPlacement (simpl_placer.cc:463-539) and refine (:541-557),
        # resident on the device: the lower bound stays in the context's
        # positions, the upper bound in its anchors; one small state record
        # comes back per outer iteration (HPWL, overflow, Convergence).
        trace = []
        ctx.set_positions(lx, ly)
        ctx.loop_reset(o.gapReferenceIter, o.gapTightFrac, o.gapRelaxedFrac, o.stallIters)
        have_upper = False
        for it in range(o.maxIters):
            self.stats["global_iters"] = it + 1
            self._legalize_into_anchors()
            have_upper = True
            st = ctx.loop_update(it + 1)
            trace.append((st["lower_hpwl"], st["upper_hpwl"], st["overflow_upper"]))
            if self.progress:
                self.progress(f"iter {it + 1}: lower {st['lower_hpwl']:.6e} upper {st['upper_hpwl']:.6e} "
                              f"lal {self.stats['spread_seconds']:.2f} s total, "
                              f"{ctx.lal_batches()} batches")
            if st["converged"]:
                break
            alpha = o.alphaBase * (1.0 + it)
            t0 = time.perf_counter()
            ctx.build_b2b_resident(alpha)
            t1 = time.perf_counter()
            self._solve_resident()
            self.stats["build_seconds"] += t1 - t0
            self.stats["solve_seconds"] += time.perf_counter() - t1
        if have_upper and o.refineIters > 0:
            done = self.stats["global_iters"]
            ctx.build_grid(ctx.bins * o.refineGridFactor, o.densityLimit)
            for r in range(o.refineIters):
                alpha = o.alphaBase * (1.0 + done + r)
                ctx.build_b2b_resident(alpha)
                self._solve_resident()
                t0 = time.perf_counter()
                self._legalize_into_anchors()
                if self.progress:
                    self.progress(f"refine {r + 1}: lal {time.perf_counter() - t0:.2f} s, "
                                  f"{ctx.lal_batches()} batches")
        if not have_upper:
            ctx.anchors_from_positions()  # upper = lower
        st = ctx.loop_update(0)
        self.stats["hpwl_final"] = st["upper_hpwl"]
        self.stats["hpwl_lower"] = st["lower_hpwl"]
        self.stats["gap"] = st["upper_hpwl"] - st["lower_hpwl"]
        self.stats["overflow_lower"] = st["overflow_lower"]
        self.stats["overflow_final"] = st["overflow_upper"]
        ux, uy = ctx.get_anchors()
        lx, ly = ctx.get_positions()
        return PlaceResult(ux, uy, self.stats, trace, lx, ly)


def main(argv=None) -> int:
    """Global placement of a Bookshelf design (or a synthetic workload) on the
    GPU; prints the statistics as JSON and writes the upper-bound placement as
    a Bookshelf .pl (src/kt_flowMgr.cc:508-521).

        python -m ktplace_amd.placer path/to/ibm01 --out placed.pl
        python -m ktplace_amd.placer --workload cells210k --bins 512
    """
    import argparse
    import json
    import sys

    from . import design as dz

    ap = argparse.ArgumentParser(description="SimPL global placement on the GPU")
    ap.add_argument("design", nargs="?", help="Bookshelf design directory (<dir>/<name>.nodes ...)")
    ap.add_argument("--workload", help="synthetic workload name instead of a design directory")
    ap.add_argument("--out", help="write the placement here as a Bookshelf .pl (design directories only)")
    ap.add_argument("--bins", type=int, default=0, help="bins per axis (0: the reference's default rule)")
    ap.add_argument("--init-net", default="star", choices=("star", "b2b", "star_b2b"))
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "exact"),
                    help="exact: the reference's operation order in the solver (bit-identical runs)")
    args = ap.parse_args(argv)
    if bool(args.design) == bool(args.workload):
        ap.error("give a design directory or --workload")
    d = dz.read_bookshelf(args.design) if args.design else dz.workload(args.workload)
    prec = simpl.KT_PREC_FP64_EXACT if args.precision == "exact" else simpl.KT_PREC_FP64
    opts = SimplOptions(binsPerAxis=args.bins, initialNetModel=args.init_net, seed=args.seed, precision=prec)
    t0 = time.perf_counter()
    res = Placer(d, opts, progress=lambda m: print(m, file=sys.stderr, flush=True)).run()
    out = dict(design=d.name, cells=d.n, nets=d.flat.num_nets, seconds=round(time.perf_counter() - t0, 3),
               **{k: (float(v) if isinstance(v, (float, np.floating)) else v) for k, v in res.stats.items()})
    if args.out:
        if not args.design:
            ap.error("--out needs a design directory (a synthetic design has no names)")
        dz.write_pl(d, args.out, res.x, res.y)
        out["placement"] = args.out
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
