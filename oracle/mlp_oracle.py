"""Test infrastructure, not product: the CPU statement of the learned face closure's network, against which the
HIP kernels (wadi2d_amd/csrc/mlp.hip, mlp.hpp) are checked. Only tests/ may import it.

The network is 16 -> 64 -> 64 -> 2 with ReLU. The device evaluates it with BF16 weights, BF16 inputs and BF16
activations between layers, FP32 accumulation and FP32 biases; this module does the same arithmetic in float64
with the same rounding points (``bf16=True``), so the device's results differ from it only by FP32 accumulation
order. With ``bf16=False`` it is the plain float64 network, whose gradient is pinned against central finite
differences (tests/test_closure.py). The reference has no learned component; nothing here follows a reference
file.

Parameter vector layout (the device's): W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
"""

from __future__ import annotations

import numpy as np

IN, HID, OUT = 16, 64, 2
SHAPES = (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)), ("b3", (OUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value, ties to even (as the device's conversion), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off : off + n], np.float64).reshape(shape)
        off += n
    return out


def forward(p: np.ndarray, x: np.ndarray, bf16: bool = True):
    """y [n][2] and the hidden activations (post-ReLU) h1, h2 [n][64]; x [n][16]."""
    w = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    x = r(x)
    h1 = r(np.maximum(x @ r(w["W1"]).T + w["b1"], 0.0))
    h2 = r(np.maximum(h1 @ r(w["W2"]).T + w["b2"], 0.0))
    y = h2 @ r(w["W3"]).T + w["b3"]
    return y, h1, h2, x


def gradient(p: np.ndarray, x: np.ndarray, t: np.ndarray, mask: np.ndarray, norm: float, bf16: bool = True):
    """Loss sum(mask * (y - t)^2) / norm and its gradient with respect to p, through the forward pass above (the
    rounding treated as the identity, as on the device)."""
    w = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    y, h1, h2, xr = forward(p, x, bf16)
    e = y - t
    loss = float((mask * e * e).sum() / norm)
    dy = 2.0 * mask * e / norm
    dz2 = (dy @ r(w["W3"])) * (h2 > 0)
    dz1 = (dz2 @ r(w["W2"])) * (h1 > 0)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0), "W2": dz2.T @ h1, "b2": dz2.sum(0), "W1": dz1.T @ xr, "b1": dz1.sum(0)}
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def init(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-bound, bound, shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts)
