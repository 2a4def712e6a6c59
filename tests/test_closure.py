"""The learned face closure. Its network's CPU statement (oracle/mlp_oracle.py) is pinned by central finite
differences and by the BF16 rounding rule; the HIP kernels (csrc/mlp.hip: matrix-core forward, FP32 backward,
Adam) are held to it on the GPU, and training on the GPU must fit a known function."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as ora


def _batch(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 16))
    t = rng.normal(size=(n, 2))
    mask = (rng.random((n, 2)) > 0.25).astype(np.float64)
    return x, t, mask


def test_oracle_gradient_is_the_finite_difference_of_its_loss():
    p = ora.init(1)
    x, t, mask = _batch(29)
    loss, g = ora.gradient(p, x, t, mask, 29.0, bf16=False)
    fd = np.zeros_like(p)
    eps = 1e-6
    for i in range(p.size):
        q = p.copy()
        q[i] += eps
        a, _ = ora.gradient(q, x, t, mask, 29.0, bf16=False)
        q[i] -= 2 * eps
        b, _ = ora.gradient(q, x, t, mask, 29.0, bf16=False)
        fd[i] = (a - b) / (2 * eps)
    assert np.abs(fd - g).max() <= 1e-7 * np.abs(g).max()
    assert loss > 0


def test_bf16_rounding_is_nearest_even_and_matches_torch():
    import torch

    one = 1.0
    ulp = 2.0**-7  # BF16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3.0e-3, 1e30])
    want = np.array([one, one, one + 2 * ulp, one, -2.5, None, None], dtype=object)
    got = ora.bf16_round(vals)
    for g, w in zip(got[:5], want[:5], strict=True):
        assert g == w
    ref = torch.tensor(vals, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(got, ref)
    rng = np.random.default_rng(3)
    r = rng.normal(size=10000) * 10.0 ** rng.integers(-6, 6, 10000)
    assert np.array_equal(ora.bf16_round(r), torch.tensor(r, dtype=torch.float32).to(torch.bfloat16).double().numpy())


def test_a_dead_relu_layer_gives_the_bias():
    """All first-layer pre-activations negative: the output is W3 relu(b2) + b3 whatever the input."""
    p = ora.init(2)
    w = ora.unpack(p)
    p = p.copy()
    p[: 16 * 64] = -np.abs(p[: 16 * 64])
    p[16 * 64 : 16 * 64 + 64] = -100.0
    b2 = np.linspace(-1, 1, 64)
    off = 16 * 64 + 64 + 64 * 64
    p[off : off + 64] = b2
    x = np.abs(_batch(5)[0])
    y, h1, _, _ = ora.forward(p, x)
    assert not h1.any()
    want = ora.bf16_round(np.maximum(b2, 0)) @ ora.bf16_round(w["W3"]).T + w["b3"]
    assert np.allclose(y, want[None, :])


@pytest.mark.gpu
def test_matrix_core_forward_matches_the_oracle():
    from wadi2d_amd.closure.net import Batch, Net

    p = ora.init(4).astype(np.float32)
    x, t, mask = _batch(1000, 5)
    net = Net(p)
    y = net.forward(Batch(x, t, mask)).astype(np.float64)
    want, _, _, _ = ora.forward(p.astype(np.float64), x)
    err = np.abs(y - want).max() / np.abs(want).max()
    print(f"forward max rel error {err:.2e}", flush=True)
    assert err <= 2e-3  # FP32 accumulation order; a wrong lane layout is off by O(1)


@pytest.mark.gpu
def test_device_gradient_matches_the_oracle():
    from wadi2d_amd.closure.net import Batch, Net

    p = ora.init(6).astype(np.float32)
    x, t, mask = _batch(3000, 7)
    net = Net(p)
    loss = net.gradient(Batch(x, t, mask))
    g = net.gradient_host()[: ora.NPARAM].astype(np.float64)
    wl, wg = ora.gradient(p.astype(np.float64), x, t, mask, 3000.0)
    err = np.abs(g - wg).max() / np.abs(wg).max()
    print(f"loss {loss:.6f}/{wl:.6f}, gradient max rel error {err:.2e}", flush=True)
    assert loss == pytest.approx(wl, rel=1e-4)
    assert err <= 5e-3


@pytest.mark.gpu
def test_training_on_the_gpu_fits_a_smooth_function():
    from wadi2d_amd.closure.net import Batch, Net

    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (20000, 16))
    t = np.stack((np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.maximum(x[:, 3], 0) - 0.3 * x[:, 4]), axis=1)
    data = Batch(x, t, np.ones_like(t))
    mini = Batch.empty(1024)
    net = Net(seed=9)
    first = None
    for _ in range(600):
        data.take(rng.integers(0, x.shape[0], 1024), mini)
        loss = net.gradient(mini)
        net.adam(lr=3e-3)
        first = loss if first is None else first
    full = net.gradient(data)
    print(f"loss {first:.4f} -> {full:.4f}", flush=True)
    assert full < 0.05 * first
