"""Launch-shape autotuner for the two-axis Jacobi-PCG kernels.

Chooses the SpMV layout (lane groups per row, or entry blocks), the lanes
per row (threads cooperating on one CSR row: 1, 2, 4, 8 or 16), the tile
(rows per 1024-thread block), rows in flight, the XCD split and the LDS
window (`halo`: entries of p staged on each side of a tile) by timing a fixed number of CG iterations of the context's current
systems on the GPU, from the same saved start each time.  First lanes,
tile, rows and XCD split without staging; then, with the best rows and XCD
split, every lanes, tile and halo whose window fits the LDS budget.  The search
costs a few hundred iterations, i.e. a fraction of one solve; results are
cached per (rows, nnz, precision) signature for the life of the process.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

from . import simpl

LANES = (4, 8, 16)
TILES = (1024, 2048, 4096)
ROWS = (2, 4, 8)
XCD = (1, 0)
HALOS = (2048, 4096, 6144, 8192, 12288, 16384)
HALO_TILES = (1024, 2048, 4096)
FLAT_ENTRIES = (4096, 8192)  # entry-block SpMV: entries per block
FLAT_UNROLL = (4, 8)         # entry-block SpMV: entries per thread and round
FLAT_ROW_LANES = (1, 4)      # entry-block SpMV: lanes per row in the row sums
FLAT_TILES = (2048, 4096)    # tiles (of the vector kernels) tried with entry blocks
WINDOW_BYTES = 156 * 1024  # LDS a staged SpMV block may take

_cache: dict = {}


@dataclass
class TuneResult:
    lanes: int
    tile: int
    ms_per_iteration: float
    table: list
    xcd_split: int = 1
    rows: int = 2
    halo: int = 0
    layout: int = 0
    block_entries: int = 0
    unroll: int = 0
    row_lanes: int = 0


def _signature(ctx: "simpl.Context", precision: int):
    rx, nx = ctx.system_size(0)
    ry, ny = ctx.system_size(1)
    return (rx, nx, ry, ny, precision)


def window_fits(tile: int, halo: int, precision: int) -> bool:
    width = 4 if precision == simpl.KT_PREC_MIXED else 8
    return (tile + 2 * halo) * width <= WINDOW_BYTES


def autotune(ctx: "simpl.Context", precision: int = simpl.KT_PREC_FP64, iterations: int = 48,
             lanes=LANES, tiles=TILES, use_graphs: bool = True, force: bool = False,
             halos=HALOS) -> TuneResult:
    """Time `iterations` CG iterations per candidate and apply the fastest.

    The context's positions are saved before and restored after the search,
    so the caller's state is unchanged.  `halos=()` leaves staging off.
    """
    key = _signature(ctx, precision)
    if not force and key in _cache:
        best = _cache[key]
        ctx.set_tuning(lanes=best.lanes, tile=best.tile, use_graphs=use_graphs,
                       xcd_split=best.xcd_split, rows=best.rows, halo=best.halo, layout=best.layout,
                       block_entries=best.block_entries, unroll=best.unroll, row_lanes=best.row_lanes)
        return best
    rows = key[0]
    ctx.save_positions()
    table = []

    def timed(g, tile, xs, rr, halo, layout=0, entries=0, unroll=0, row_lanes=0):
        ctx.set_tuning(lanes=g, tile=tile, use_graphs=use_graphs, xcd_split=xs, rows=rr, halo=halo,
                       layout=layout, block_entries=entries, unroll=unroll, row_lanes=row_lanes)
        ctx.restore_positions()
        ctx.solve_resident(0.0, 2, precision=precision)  # graph capture, first touch
        ctx.restore_positions()
        ctx.sync()
        t0 = time.perf_counter()
        r = ctx.solve_resident(0.0, iterations, precision=precision, check_every=iterations)
        ctx.sync()
        dt = time.perf_counter() - t0
        its = max(r["iterations"], 1)
        table.append((g, tile, xs, rr, halo, layout, entries, unroll, row_lanes, 1e3 * dt / its))

    small = lambda t: not (t > 4 * max(rows, 1024) and t != tiles[0])
    for g, tile, xs, rr in [(g, t, x, rr) for t in tiles for g in lanes for x in XCD for rr in ROWS if small(t)]:
        timed(g, tile, xs, rr, 0)
    _, _, xs, rr, _, _, _, _, _, _ = min(table, key=lambda t: t[-1])
    for tile in HALO_TILES:
        for halo in halos:
            if window_fits(tile, halo, precision) and halo < rows and small(tile):
                for g in lanes:
                    timed(g, tile, xs, rr, halo)
    # entry-block SpMV (lanes, rows in flight and halo do not apply; the tile
    # and XCD split shape the vector kernels)
    for tile in FLAT_TILES:
        for x in XCD:
            if small(tile):
                for entries in FLAT_ENTRIES:
                    for unroll in FLAT_UNROLL:
                        for row_lanes in FLAT_ROW_LANES:
                            timed(4, tile, x, 2, 0, simpl.KT_SPMV_ENTRY_BLOCKS, entries, unroll, row_lanes)
    # the closest candidates again, over twice the iterations, twice each:
    # single short timings of near-equal shapes are noisy
    finalists = sorted(table, key=lambda t: t[-1])[:4]
    table = [t for t in table if t not in finalists]
    for cand in finalists:
        best_ms = float("inf")
        for _ in range(2):
            ctx.set_tuning(lanes=cand[0], tile=cand[1], use_graphs=use_graphs, xcd_split=cand[2], rows=cand[3],
                           halo=cand[4], layout=cand[5], block_entries=cand[6], unroll=cand[7], row_lanes=cand[8])
            ctx.restore_positions()
            ctx.sync()
            t0 = time.perf_counter()
            r = ctx.solve_resident(0.0, 2 * iterations, precision=precision, check_every=2 * iterations)
            ctx.sync()
            best_ms = min(best_ms, 1e3 * (time.perf_counter() - t0) / max(r["iterations"], 1))
        table.append((*cand[:9], best_ms))
    ctx.restore_positions()
    g, tile, xs, rr, halo, layout, entries, unroll, row_lanes, ms = min(table, key=lambda t: t[-1])
    best = TuneResult(lanes=g, tile=tile, ms_per_iteration=ms, table=table, xcd_split=xs, rows=rr, halo=halo,
                      layout=layout, block_entries=entries, unroll=unroll, row_lanes=row_lanes)
    _cache[key] = best
    ctx.set_tuning(lanes=g, tile=tile, use_graphs=use_graphs, xcd_split=xs, rows=rr, halo=halo, layout=layout,
                   block_entries=entries, unroll=unroll, row_lanes=row_lanes)
    return best
