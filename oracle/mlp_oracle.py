"""CPU oracle of the warm-start network (test infrastructure only).

Only tests/ may import this module; the product path never does.  The network
has no counterpart in the reference: this is the plain FP64 mathematics of the
3-layer MLP the GPU trains and evaluates (32 inputs, ReLU hidden layers of 128,
2 outputs) and of its loss

    L = (1/B) sum_s w_s |net(f_s) - t_s|^2,

with the parameters packed as the C ABI packs them (W1 [128][32], b1 [128],
W2 [128][128], b2 [128], W3 [2][128], b3 [2], output unit first).
`device_params` rounds the weight matrices to BF16 (nearest even) as the GPU
kernels use them.

Pinned by tests/test_mlp_oracle.py: the forward pass and the gradient against
torch's float64 autograd (an independent implementation), and the gradient
against central finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, FEATURES = 32, 128, 2, 17
SHAPES = (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (OUT, HID)), ("b3", (OUT,)))
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[off:off + k], np.float64).reshape(shape)
        off += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(a) -> np.ndarray:
    """Round float32 values to BF16, nearest even (returned as float64)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def device_params(p: np.ndarray) -> np.ndarray:
    """The parameters as the kernels see them: weight matrices in BF16."""
    d = unpack(np.asarray(p, np.float32).astype(np.float64))
    for k in ("W1", "W2", "W3"):
        d[k] = bf16(d[k])
    return pack(d)


def pad_inputs(features: np.ndarray) -> np.ndarray:
    f = np.asarray(features, np.float64)
    x = np.zeros((f.shape[0], IN))
    x[:, :f.shape[1]] = f
    return x


def forward(p: np.ndarray, features: np.ndarray):
    d = unpack(p)
    x = pad_inputs(features)
    z1 = x @ d["W1"].T + d["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(z2, 0.0)
    out = h2 @ d["W3"].T + d["b3"]
    return out, (x, z1, h1, z2, h2)


def loss_and_gradient(p: np.ndarray, features: np.ndarray, targets: np.ndarray, weights=None):
    """L and dL/dp (packed like p)."""
    d = unpack(p)
    out, (x, z1, h1, z2, h2) = forward(p, features)
    t = np.asarray(targets, np.float64)
    w = np.ones(len(t)) if weights is None else np.asarray(weights, np.float64)
    B = len(t)
    e = out - t
    loss = float(np.sum(w[:, None] * e * e) / B)
    dout = 2.0 * w[:, None] * e / B
    g = {"W3": dout.T @ h2, "b3": dout.sum(0)}
    dz2 = (dout @ d["W3"]) * (z2 > 0)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ d["W2"]) * (z1 > 0)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return loss, pack(g)

