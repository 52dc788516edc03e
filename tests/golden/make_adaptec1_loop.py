"""Writes adaptec1_loop.json: the oracle's full SimPL run (reference defaults:
star initial rounds, default bin grid, refinement) on the ISPD 2005 adaptec1
design, read by the library's Bookshelf reader from
tests/golden/bookshelf/adaptec1 (the benchmark files as the reference's
benchmark/ directory holds them), for the GPU loop's comparison.  Also records
the design's sizes as the reader sees them (the reference's CI checks the same
design's load line: 211447 cells, 543 terminals, 221142 nets, 944053 pins,
890 rows) and the final HPWL and overflow of the same run from seeds 1-4.

    python tests/golden/make_adaptec1_loop.py
"""
import json
import os
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

import oracle_lib as ol  # noqa: E402
from ktplace_amd import design as dz  # noqa: E402


def main():
    d = dz.read_bookshelf(HERE / "bookshelf" / "adaptec1")
    o = ol.Oracle.of(d)
    ol.set_threads(os.cpu_count() or 1)
    t = time.perf_counter()
    x, y, st, trace = o.place(initial_net_model=0)
    out = dict(design="adaptec1", source="ISPD 2005 adaptec1 (Bookshelf)", initial_net_model="star",
               bins_per_axis=0, seconds=time.perf_counter() - t,
               sizes=dict(vertices=d.flat.num_vertices, movable=d.n, nets=d.flat.num_nets,
                          pins=d.flat.num_pins, pins_in_file=int(len(d.raw.pin_vertex)),
                          rows=int(len(d.raw.row_y)), die=[float(v) for v in d.die]),
               stats={k: float(v) for k, v in st.items()},
               trace=[[float(v) for v in row] for row in trace])
    print(json.dumps(out["stats"]), flush=True)
    spread = {}
    for seed in range(1, 5):
        s = o.place(initial_net_model=0, seed=seed)[2]
        spread[str(seed)] = dict(hpwl_final=float(s["hpwl_final"]), overflow_final=float(s["overflow_final"]))
        print(seed, spread[str(seed)], flush=True)
    out["seed_spread"] = spread
    (HERE / "adaptec1_loop.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
