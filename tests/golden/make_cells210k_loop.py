"""Writes cells210k_loop.json: the oracle's full SimPL run on the 210,000-cell
workload (seed 1; star initial rounds, 512 x 512 bins, reference defaults
otherwise), for the GPU loop's HPWL comparison.  Takes several minutes of CPU.

    python tests/golden/make_cells210k_loop.py
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
    d = dz.workload("cells210k")
    o = ol.Oracle.of(d)
    ol.set_threads(os.cpu_count() or 1)
    t = time.perf_counter()
    x, y, st, trace = o.place(initial_net_model=0, bins_per_axis=512)
    out = dict(workload="cells210k", seed=1, initial_net_model="star", bins_per_axis=512,
               seconds=time.perf_counter() - t, stats={k: float(v) for k, v in st.items()},
               trace=[[float(v) for v in row] for row in trace])
    (HERE / "cells210k_loop.json").write_text(json.dumps(out, indent=1))
    print(json.dumps(out["stats"]))


if __name__ == "__main__":
    main()
