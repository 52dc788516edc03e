"""The CPU oracle of the warm-start network, pinned against torch's float64
autograd (an independent implementation) and central finite differences."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402


def _random_case(seed, B=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mo.PARAMS)
    f = rng.normal(0, 1, (B, mo.FEATURES))
    t = rng.normal(0, 1, (B, 2))
    w = rng.uniform(0.2, 2.0, B)
    return p, f, t, w


def test_layout_size():
    assert mo.PARAMS == 20994
    d = mo.unpack(np.arange(mo.PARAMS, dtype=np.float64))
    assert d["W1"][1, 0] == 32 and d["b1"][0] == 4096 and d["W2"][0, 0] == 4224
    assert d["W3"][1, 127] == 20991 and d["b3"][1] == 20993
    assert np.array_equal(mo.pack(d), np.arange(mo.PARAMS))


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_and_gradient_match_torch_autograd(seed):
    torch = pytest.importorskip("torch")
    p, f, t, w = _random_case(seed)
    loss, g = mo.loss_and_gradient(p, f, t, w)
    out, _ = mo.forward(p, f)
    d = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in mo.unpack(p).items()}
    x = torch.tensor(mo.pad_inputs(f))
    h1 = torch.relu(x @ d["W1"].T + d["b1"])
    h2 = torch.relu(h1 @ d["W2"].T + d["b2"])
    o = h2 @ d["W3"].T + d["b3"]
    tl = (torch.tensor(w)[:, None] * (o - torch.tensor(t)) ** 2).sum() / len(t)
    tl.backward()
    assert np.allclose(out, o.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert loss == pytest.approx(tl.item(), rel=1e-12)
    tg = mo.pack({k: v.grad.numpy() for k, v in d.items()})
    assert np.allclose(g, tg, rtol=1e-10, atol=1e-12)


def test_gradient_matches_finite_differences():
    p, f, t, w = _random_case(3, B=12)
    _, g = mo.loss_and_gradient(p, f, t, w)
    rng = np.random.default_rng(4)
    # entries of every parameter block, including the unused input columns
    # of W1 (their gradient is exactly zero)
    idx = np.concatenate([rng.choice(np.arange(a, b), min(6, b - a), replace=False)
                          for a, b in ((0, 4096), (4096, 4224), (4224, 20608), (20608, 20736),
                                       (20736, 20992), (20992, 20994))] + [[17, 31]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mo.loss_and_gradient(p + e, f, t, w)
        lm, _ = mo.loss_and_gradient(p - e, f, t, w)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-8), i
    assert g[17] == 0.0 and g[31] == 0.0


def test_bf16_rounding_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.00390625, 3.0e-39], np.float32)
    r = mo.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even 1.0;
    # 1 + 3 * 2^-8 is a tie rounding up to 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    normal = np.abs(vals) > 1e-37
    assert np.all(np.abs(r - vals)[normal] <= np.abs(vals[normal]) * 2 ** -8)
    assert abs(r[5] - vals[5]) <= 2.0 ** -133  # subnormal: half a BF16 step
