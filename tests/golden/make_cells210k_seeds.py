"""Writes cells210k_seeds.json: the oracle's full SimPL run on the
210,000-cell workload (star initial rounds, 512 x 512 bins, reference
defaults otherwise) from several start seeds (seedUniform's seed,
simpl_placer.cc:226-237), so the GPU loop's final HPWL can be judged
against the spread that the seed alone causes.  Takes several minutes of
CPU per seed.

    python tests/golden/make_cells210k_seeds.py [seed ...]
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

SEEDS = (20240607, 1, 2, 3, 4)


def main(seeds):
    d = dz.workload("cells210k")
    ol.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    path = HERE / "cells210k_seeds.json"
    out = json.loads(path.read_text()) if path.exists() else dict(
        workload="cells210k", initial_net_model="star", bins_per_axis=512, runs={})
    for seed in seeds:
        o = ol.Oracle.of(d)
        t = time.perf_counter()
        _, _, st, _ = o.place(initial_net_model=0, bins_per_axis=512, seed=seed)
        out["runs"][str(seed)] = dict(seconds=time.perf_counter() - t,
                                      stats={k: float(v) for k, v in st.items()})
        path.write_text(json.dumps(out, indent=1))
        print(seed, json.dumps(out["runs"][str(seed)]), flush=True)


if __name__ == "__main__":
    main([int(s) for s in sys.argv[1:]] or SEEDS)
