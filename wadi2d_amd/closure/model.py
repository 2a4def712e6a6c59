"""A trained closure (network parameters and feature standardisation), its training on the GPU, and its
installation on a model: which faces it corrects and their static features, in the device's own layout."""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from wadi2d_amd.closure.features import tile_index_of
from wadi2d_amd.closure.net import Batch, Net


@dataclass
class ClosureModel:
    params: np.ndarray  # FP32 network parameters
    mean: np.ndarray  # 16 feature means
    inv_std: np.ndarray  # 16 inverse deviations (0 for a constant feature)
    conveyance: bool = True
    rise_scale: float = 1.0  # calibration of the invert-rise output on training members

    @property
    def norm(self) -> np.ndarray:
        return np.concatenate([self.mean, self.inv_std, [self.rise_scale]]).astype(np.float32)

    def with_(self, **kw) -> ClosureModel:
        from dataclasses import replace

        return replace(self, **kw)

    def save(self, path: str) -> None:
        np.savez(path, params=self.params, mean=self.mean, inv_std=self.inv_std, conveyance=self.conveyance,
                 rise_scale=self.rise_scale)  # fmt: skip

    @classmethod
    def load(cls, path: str) -> ClosureModel:
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], f["mean"], f["inv_std"], bool(f["conveyance"]), float(f["rise_scale"]))


def standardise(x: np.ndarray):
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    inv = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-12), 0.0)
    return mean.astype(np.float32), inv.astype(np.float32)


def train(x: np.ndarray, target: np.ndarray, mask: np.ndarray, steps: int = 4000, batch: int = 4096,
          lr: float = 2e-3, seed: int = 0, holdout: float = 0.1, log=None):  # fmt: skip
    """Minibatch Adam on the GPU's own kernels; returns the model and the loss history (train, held-out)."""
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    order = rng.permutation(n)
    n_val = int(n * holdout)
    val, tr = order[:n_val], order[n_val:]
    mean, inv = standardise(x[tr])
    xs = (x - mean) * inv
    data = Batch(xs[tr], target[tr], mask[tr])
    check = Batch(xs[val], target[val], mask[val]) if n_val else None
    mini = Batch.empty(batch)
    net = Net(seed=seed)
    history = []
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        data.take(rng.integers(0, tr.size, batch), mini)
        loss = net.gradient(mini)
        rate = lr * (0.5 * (1.0 + np.cos(np.pi * step / steps)))  # cosine decay
        net.adam(lr=rate)
        if step % 250 == 0 or step == steps:
            held = net.gradient(check) if check is not None else float("nan")
            history.append((step, loss, held))
            if log is not None:
                log(f"[train] step {step}: minibatch loss {loss:.4f}, held-out {held:.4f}, {time.perf_counter() - t0:.1f} s")
    return ClosureModel(net.host_params(), mean, inv), history


def install(backend, model: ClosureModel, members: list) -> int:
    """Put the closure on the backend's faces. ``members``: (CoarseFaces, row offset in the model's grid) per
    member (one member at offset 0 for a plain model). Returns the number of faces."""
    grid = backend.grid
    rs, ls, axes, stats = [], [], [], []
    for cf, row0 in members:
        for axis in (1, 0):
            jj, ii = np.nonzero(cf.closure_faces(axis))
            st = cf.stat[axis]["static"][jj, ii]
            lj, li = (jj, ii - 1) if axis == 1 else (jj - 1, ii)
            rs.append(tile_index_of(grid, jj + row0, ii))
            ls.append(tile_index_of(grid, lj + row0, li))
            axes.append(np.full(jj.size, axis, np.int32))
            stats.append(st)
    r = np.concatenate(rs) if rs else np.zeros(0, np.int64)
    backend.closure_set(r, np.concatenate(ls) if ls else r, np.concatenate(axes) if axes else r,
                        np.concatenate(stats) if stats else np.zeros((0, 11)), model.norm, model.params,
                        model.conveyance)  # fmt: skip
    return int(r.size)

