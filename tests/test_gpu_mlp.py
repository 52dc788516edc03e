"""The warm-start network on the GPU against its FP64 CPU oracle.

The kernels multiply BF16 weights and BF16 activations with FP32
accumulation; the oracle computes in FP64 with the same BF16-rounded weights
and inputs.  Bars: outputs within 1 % relative L2 error (the hidden
activations are rounded to BF16 between layers: 2^-9 relative each);
gradients within 3 % relative L2 error per parameter block and cosine
similarity above 0.999; the loss within 1 %.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu

BLOCKS = (("W1", 0, 4096), ("b1", 4096, 4224), ("W2", 4224, 20608), ("b2", 20608, 20736),
          ("W3", 20736, 20992), ("b3", 20992, 20994))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def case(seed, n=3000):
    from ktplace_amd.learned import init_params
    rng = np.random.default_rng(seed)
    p = init_params(seed, out_scale=1.0)
    p[4096:4224] = rng.normal(0, 0.1, 128)   # non-zero biases
    p[20608:20736] = rng.normal(0, 0.1, 128)
    p[20992:] = rng.normal(0, 0.1, 2)
    f = rng.normal(0, 1, (n, mo.FEATURES)).astype(np.float32)
    t = rng.normal(0, 1, (n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return p.astype(np.float32), f, t, w


@pytest.fixture(scope="module")
def mlp():
    from ktplace_amd.learned import Mlp
    m = Mlp(0)
    yield m
    m.close()


def test_forward_matches_oracle(mlp):
    p, f, _, _ = case(1)
    mlp.set_params(p)
    out = mlp.forward(f)
    ref, _ = mo.forward(mo.device_params(p), mo.bf16(f))
    print("forward rel error", rel(out, ref), flush=True)
    assert rel(out, ref) < 1e-2
    assert np.abs(out - ref).max() < 0.05 * np.sqrt(np.mean(ref ** 2))
    # ragged count (not a multiple of 16) and a single sample
    assert np.allclose(mlp.forward(f[:37]), out[:37], rtol=0, atol=0)
    assert np.allclose(mlp.forward(f[5:6]), out[5:6], rtol=0, atol=0)


@pytest.mark.parametrize("weighted", [False, True])
def test_gradient_matches_oracle(mlp, weighted):
    p, f, t, w = case(2)
    mlp.set_params(p)
    mlp.set_data(f, t, w if weighted else None)
    samples = np.random.default_rng(3).choice(len(f), 1000, replace=False).astype(np.uint32)
    g, loss = mlp.gradient(samples)
    rl, rg = mo.loss_and_gradient(mo.device_params(p), mo.bf16(f[samples]), t[samples],
                                  w[samples] if weighted else None)
    print("loss", loss, rl, flush=True)
    assert loss == pytest.approx(rl, rel=1e-2)
    for name, a, b in BLOCKS:
        cos = float(np.dot(g[a:b], rg[a:b]) / (np.linalg.norm(g[a:b]) * np.linalg.norm(rg[a:b])))
        print(name, "rel", rel(g[a:b], rg[a:b]), "cos", cos, flush=True)
        assert rel(g[a:b], rg[a:b]) < 3e-2, name
        assert cos > 0.999, name
    # the always-zero inputs get exactly zero gradient
    W1 = g[:4096].reshape(128, 32)
    assert np.all(W1[:, mo.FEATURES:] == 0.0)


def test_gradient_is_deterministic(mlp):
    p, f, t, w = case(4)
    mlp.set_params(p)
    mlp.set_data(f, t, w)
    s = np.arange(0, 2000, 3, dtype=np.uint32)
    g1, l1 = mlp.gradient(s)
    g2, l2 = mlp.gradient(s)
    assert np.array_equal(g1, g2) and l1 == l2


def test_training_fits_a_smooth_map(mlp):
    """Adam on a known smooth regression: the loss falls by more than 20x and
    the trained network generalises to fresh inputs."""
    from ktplace_amd.learned import init_params
    rng = np.random.default_rng(5)

    def target(f):
        return np.stack([0.5 * np.sin(f[:, 0]) + 0.3 * f[:, 1], 0.2 * f[:, 0] * f[:, 2] - 0.1 * f[:, 3]], 1)

    f = rng.normal(0, 1, (20000, mo.FEATURES)).astype(np.float32)
    t = target(f).astype(np.float32)
    mlp.set_params(init_params(6))
    mlp.set_data(f, t)
    loss = mlp.train(1500, 1024, 2e-3, 7)
    print("loss", loss[:3], loss[-3:], flush=True)
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < loss[:5].mean() / 20
    ft = rng.normal(0, 1, (2000, mo.FEATURES)).astype(np.float32)
    err = np.mean(np.sum((mlp.forward(ft) - target(ft)) ** 2, 1))
    assert err < 0.05 * np.mean(np.sum(target(ft) ** 2, 1))
    # the trainer's master parameters evaluate the same on the CPU oracle
    ref, _ = mo.forward(mo.device_params(mlp.get_params()), mo.bf16(ft))
    assert rel(mlp.forward(ft), ref) < 1e-2


def test_rejects_bad_input(mlp):
    from ktplace_amd.simpl import KtError
    with pytest.raises(KtError):
        mlp.set_params(np.zeros(10, np.float32))
    bad = np.zeros(20994, np.float32)
    bad[3] = np.nan
    with pytest.raises(KtError):
        mlp.set_params(bad)
    mlp.set_params(np.zeros(20994, np.float32))
    mlp.set_data(np.zeros((64, mo.FEATURES), np.float32), np.zeros((64, 2), np.float32))
    with pytest.raises(KtError):
        mlp.train(1, 33, 1e-3, 0)   # batch not a multiple of 32
    with pytest.raises(KtError):
        mlp.gradient(np.array([64], np.uint32))  # sample out of range
