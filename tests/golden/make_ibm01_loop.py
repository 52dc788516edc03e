"""Writes ibm01_loop.json: the oracle's full SimPL run (reference defaults:
star initial rounds, default bin grid, refinement) on the ICCAD04 ibm01
design, read by the library's Bookshelf reader from tests/golden/bookshelf/ibm01
(the benchmark files as the reference's benchmark/ directory holds them),
for the GPU loop's HPWL comparison.  Also records the design's sizes as the
reader sees them, and the final HPWL of the same run from seeds 1-7 (the
spread the algorithm itself shows on this design: a run whose arithmetic
differs by rounding, like the GPU's default solver, follows another
trajectory through the legalizer and lands within it).

    python tests/golden/make_ibm01_loop.py
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
    d = dz.read_bookshelf(HERE / "bookshelf" / "ibm01")
    o = ol.Oracle.of(d)
    ol.set_threads(os.cpu_count() or 1)
    t = time.perf_counter()
    x, y, st, trace = o.place(initial_net_model=0)
    out = dict(design="ibm01", source="ICCAD04 ibm01 (Bookshelf)", initial_net_model="star",
               bins_per_axis=0, seconds=time.perf_counter() - t,
               sizes=dict(vertices=d.flat.num_vertices, movable=d.n, nets=d.flat.num_nets,
                          pins=d.flat.num_pins, rows=int(len(d.raw.row_y)),
                          die=[float(v) for v in d.die]),
               stats={k: float(v) for k, v in st.items()},
               trace=[[float(v) for v in row] for row in trace])
    out["seed_spread"] = {str(seed): float(o.place(initial_net_model=0, seed=seed)[2]["hpwl_final"])
                          for seed in range(1, 8)}
    (HERE / "ibm01_loop.json").write_text(json.dumps(out, indent=1))
    print(json.dumps(out["stats"]))


if __name__ == "__main__":
    main()
