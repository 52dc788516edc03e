"""kt_hpwl_noalloc (the HPWL entry point that allocates nothing) against
kt_hpwl through the C ABI: bit-identical results on two calls in a row, the
second on a smaller netlist (both contexts share the GPU's fixed scratch),
then the larger one again; the error for missing positions."""
import numpy as np
import pytest

from ktplace_amd import design as dz
from ktplace_amd import simpl

pytestmark = pytest.mark.gpu


def clumped(d, seed):
    rng = np.random.default_rng(seed)
    W, H = d.die[2], d.die[3]
    return (np.clip(rng.normal(W / 2, W / 6, d.n), 0, W - d.flat.width),
            np.clip(rng.normal(H / 2, H / 6, d.n), 0, H - d.flat.height))


def test_noalloc_hpwl_matches_kt_hpwl_on_two_calls():
    big = dz.workload("cells210k")       # 220,000 nets: 54 chunks
    small = dz.generate(3000, 3300, 60, seed=4)  # 3,300 nets: 1 chunk
    with simpl.Context(0) as cb, simpl.Context(0) as cs:
        cb.upload(big)
        cs.upload(small)
        xb, yb = clumped(big, 1)
        xs, ys = clumped(small, 2)
        cb.set_positions(xb, yb)
        cs.set_positions(xs, ys)
        first = cb.hpwl_noalloc()
        second = cs.hpwl_noalloc()   # smaller input, same scratch
        again = cb.hpwl_noalloc()
        print(f"no-alloc HPWL {first!r} {second!r} {again!r}", flush=True)
        assert first == cb.hpwl() and again == first
        assert second == cs.hpwl()
        # a new placement of the small design changes the answer the same way
        xs2, ys2 = clumped(small, 3)
        cs.set_positions(xs2, ys2)
        assert cs.hpwl_noalloc() == cs.hpwl() != second


def test_noalloc_hpwl_needs_positions():
    d = dz.generate(3000, 3300, 60, seed=4)
    with simpl.Context(0) as c:
        c.upload(d)
        with pytest.raises(simpl.KtError, match="positions"):
            c.hpwl_noalloc()
