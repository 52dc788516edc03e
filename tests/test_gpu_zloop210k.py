"""The full SimPL loop on the 210,000-cell workload (540 macros, 512 x 512
bins, star initial rounds, refinement on the 1536 x 1536 grid, the
reference's other defaults) against the oracle's run recorded in
tests/golden/cells210k_loop.json (tests/golden/make_cells210k_loop.py).

With the reference-order solver every step is bit-exact, so the GPU run ends
on the oracle's final HPWL exactly; with the default solver the trajectory
differs by rounding, and the final HPWL is judged over several start seeds
against the oracle's runs from the same seeds.  Each run takes about half a
minute on one MI355X.
"""
import json
from pathlib import Path

import pytest

from ktplace_amd import design as dz

pytestmark = pytest.mark.gpu

REF = json.loads((Path(__file__).parent / "golden" / "cells210k_loop.json").read_text())
SEEDS = json.loads((Path(__file__).parent / "golden" / "cells210k_seeds.json").read_text())


@pytest.fixture(scope="module")
def d210():
    return dz.workload("cells210k")


def run(d, ctx, precision):
    from ktplace_amd.placer import Placer, SimplOptions
    return Placer(d, SimplOptions(binsPerAxis=512, precision=precision), ctx).run()


def test_reference_order_loop_is_identical(d210, gpu_ctx_factory):
    from ktplace_amd.simpl import KT_PREC_FP64_EXACT
    ctx = gpu_ctx_factory()
    res = run(d210, ctx, KT_PREC_FP64_EXACT)
    st = REF["stats"]
    print("gpu", res.stats["hpwl_final"], "oracle", st["hpwl_final"], flush=True)
    assert res.stats["init_iters"] == st["init_iters"]
    assert res.stats["global_iters"] == st["global_iters"]
    assert res.stats["cg_iterations"] == st["cg_iterations"]
    assert res.stats["hpwl_lower"] == st["hpwl_lower"]
    assert res.stats["hpwl_final"] == st["hpwl_final"]
    ctx.close()


def test_default_solver_loop_hpwl_across_seeds(d210, gpu_ctx_factory):
    """The default solver sums in another order than the reference, and the
    loop's discrete legalization decisions turn that rounding into another
    trajectory: a run's final HPWL is one sample of a spread (the oracle's
    own runs differ by more than 1 % between start seeds), and runs of the
    same seed are not paired.  So the bar is on the distributions over the
    start seeds of tests/golden/cells210k_seeds.json (the oracle's runs,
    tests/golden/make_cells210k_seeds.py): the mean final HPWL of the GPU
    runs is within 0.5 % of the mean of the oracle's runs of the same seeds,
    and no GPU run is more than 5 % from the oracle's mean."""
    from ktplace_amd.placer import Placer, SimplOptions
    from ktplace_amd.simpl import KT_PREC_FP64
    runs = SEEDS["runs"]
    ref_mean = sum(r["stats"]["hpwl_final"] for r in runs.values()) / len(runs)
    got = []
    for seed, ref in sorted(runs.items()):
        ctx = gpu_ctx_factory()
        res = Placer(d210, SimplOptions(binsPerAxis=512, precision=KT_PREC_FP64, seed=int(seed)), ctx).run()
        ctx.close()
        got.append(res.stats["hpwl_final"])
        print(f"seed {seed}: gpu {got[-1]:.1f} oracle {ref['stats']['hpwl_final']:.1f}, "
              f"{res.stats['global_iters']} global iterations", flush=True)
        assert res.stats["init_iters"] == ref["stats"]["init_iters"]
        assert abs(got[-1] - ref_mean) <= 0.05 * ref_mean, (seed, got[-1], ref_mean)
    mean = sum(got) / len(got)
    print(f"mean final HPWL: gpu {mean:.1f}, oracle {ref_mean:.1f} ({100 * (mean - ref_mean) / ref_mean:+.3f} %) "
          f"over {len(got)} seeds", flush=True)
    assert len(got) >= 3 and abs(mean - ref_mean) <= 0.005 * ref_mean
