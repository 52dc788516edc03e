"""Launch-shape autotuner for the two-axis Jacobi-PCG kernels.

Chooses the SpMV layout (lane groups per row, or entry blocks), the lanes
per row (threads cooperating on one CSR row: 1, 2, 4, 8 or 16), the tile
(rows per 1024-thread block), rows in flight, the XCD split, the LDS
window (`halo`: entries of p staged on each side of a tile) and the
entry-block shape (threads and entries per block, lanes per row in the row
sums, 3- or 4-byte column indices, one entry or four consecutive entries per
load) by timing a fixed number of CG iterations of the context's current
systems on the GPU, from the same saved start each time.  First lanes,
tile, rows and XCD split without staging; then, with the best rows and XCD
split, every lanes, tile and halo whose window fits the LDS budget; then the
entry-block shapes; then the four fastest again over longer runs.  The search
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
FLAT_XCD = (1, 0, 2)  # entry-block SpMV: 2 = one contiguous share of the blocks per XCD
HALOS = (2048, 4096, 6144, 8192, 12288, 16384)
HALO_TILES = (1024, 2048, 4096)
# entry-block SpMV: (threads per block, entries per block) shapes
FLAT_SHAPES = ((256, 2048), (512, 4096), (1024, 8192))
FLAT_ROW_LANES = (1, 4)      # entry-block SpMV: lanes per row in the row sums
FLAT_TILES = (2048, 4096, 8192)  # tiles (of the vector kernels) tried with entry blocks
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
    block_threads: int = 0
    col_bytes: int = 0
    wide: int = 0

    def kwargs(self) -> dict:
        """The choice as `Context.set_tuning` keyword arguments."""
        return dict(lanes=self.lanes, tile=self.tile, xcd_split=self.xcd_split, rows=self.rows, halo=self.halo,
                    layout=self.layout, block_entries=self.block_entries, unroll=self.unroll,
                    row_lanes=self.row_lanes, block_threads=self.block_threads, col_bytes=self.col_bytes,
                    wide=self.wide)


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
    `table` holds (set_tuning keyword arguments, ms per iteration) pairs.
    """
    key = _signature(ctx, precision)
    if not force and key in _cache:
        best = _cache[key]
        ctx.set_tuning(use_graphs=use_graphs, **best.kwargs())
        return best
    rows = key[0]
    ctx.save_positions()
    table = []

    def run(cand: dict, its: int) -> float:
        ctx.set_tuning(use_graphs=use_graphs, **cand)
        ctx.restore_positions()
        ctx.sync()
        t0 = time.perf_counter()
        r = ctx.solve_resident(0.0, its, precision=precision, check_every=its)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / max(r["iterations"], 1)

    def timed(**cand):
        ctx.set_tuning(use_graphs=use_graphs, **cand)
        ctx.restore_positions()
        ctx.solve_resident(0.0, 2, precision=precision)  # graph capture, first touch
        table.append((cand, run(cand, iterations)))

    small = lambda t: not (t > 4 * max(rows, 1024) and t != tiles[0])
    for g, tile, xs, rr in [(g, t, x, rr) for t in tiles for g in lanes for x in XCD for rr in ROWS if small(t)]:
        timed(lanes=g, tile=tile, xcd_split=xs, rows=rr)
    best_rows = min(table, key=lambda t: t[1])[0]
    xs, rr = best_rows["xcd_split"], best_rows["rows"]
    for tile in HALO_TILES:
        for halo in halos:
            if window_fits(tile, halo, precision) and halo < rows and small(tile):
                for g in lanes:
                    timed(lanes=g, tile=tile, xcd_split=xs, rows=rr, halo=halo)
    # entry-block SpMV (lanes, rows in flight and halo do not apply; the tile
    # and XCD split shape the vector kernels)
    for tile in FLAT_TILES:
        for x in FLAT_XCD:
            if small(tile):
                for threads, entries in FLAT_SHAPES:
                    for row_lanes in FLAT_ROW_LANES:
                        for col_bytes in ((4, 3) if threads == 256 and rows < (1 << 24) else (4,)):
                            # wide loads are built for 256-thread blocks with 4 lanes per row
                            for wide in ((0, 1) if threads == 256 and row_lanes == 4 else (0,)):
                                timed(lanes=4, tile=tile, xcd_split=x, rows=2, layout=simpl.KT_SPMV_ENTRY_BLOCKS,
                                      block_entries=entries, unroll=4, row_lanes=row_lanes, block_threads=threads,
                                      col_bytes=col_bytes, wide=wide)
    # the closest candidates again, over twice the iterations, twice each:
    # single short timings of near-equal shapes are noisy
    table.sort(key=lambda t: t[1])
    finalists, table = table[:4], table[4:]
    for cand, _ in finalists:
        table.append((cand, min(run(cand, 2 * iterations) for _ in range(2))))
    ctx.restore_positions()
    cand, ms = min(table, key=lambda t: t[1])
    best = TuneResult(ms_per_iteration=ms, table=table, **{"xcd_split": 1, "rows": 2, **cand})
    _cache[key] = best
    ctx.set_tuning(use_graphs=use_graphs, **best.kwargs())
    return best
