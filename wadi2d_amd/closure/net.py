"""The closure's network on the device: FP32 master parameters, the BF16 weight block the matrix-core forward pass
reads, Adam moments; forward, gradient and Adam step through the C-ABI (csrc/mlp.hip). Device memory is the
library's own (``Buffer``); work goes to the default stream and every read back waits for it."""

from __future__ import annotations

import ctypes

import numpy as np

from wadi2d_amd.solver.hip_backend import api, check

IN, HID, OUT = 16, 64, 2


def sizes() -> tuple[int, int]:
    a, b = ctypes.c_int32(0), ctypes.c_int32(0)
    check(api().w2d_mlp_sizes(ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def he_init(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the parameter layout W1, b1, W2, b2, W3, b3."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in ((HID, IN), (HID,), (HID, HID), (HID,), (OUT, HID), (OUT,)):
        if len(shape) == 2:
            bound = np.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-bound, bound, shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """Values rounded to BF16 (nearest, ties to even), as uint16 bits."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class Buffer:
    """A device array owned by this object: shape, dtype, upload and download (synchronous)."""

    def __init__(self, shape, dtype) -> None:
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(api().w2d_dev_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def of(cls, a: np.ndarray) -> Buffer:
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes for a buffer of {self.nbytes}")
        check(api().w2d_dev_copy(self.ptr, a.ctypes.data, self.nbytes, 0))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(api().w2d_dev_copy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def __del__(self) -> None:
        p = getattr(self, "ptr", None)
        if p is not None and p.value:
            try:
                api().w2d_dev_free(p)
            except Exception:  # interpreter shutdown: the library may already be gone
                pass
            self.ptr = ctypes.c_void_p()


class Batch:
    """Training samples on the device: BF16 inputs [n][16], FP32 targets and masks [n][2]."""

    def __init__(self, x: np.ndarray, target: np.ndarray, mask: np.ndarray) -> None:
        self.n = int(x.shape[0])
        self.x = Buffer.of(bf16_bits(x).reshape(self.n, IN))
        self.t = Buffer.of(np.asarray(target, np.float32).reshape(self.n, OUT))
        self.m = Buffer.of(np.asarray(mask, np.float32).reshape(self.n, OUT))

    @classmethod
    def empty(cls, n: int) -> Batch:
        b = cls.__new__(cls)
        b.n = n
        b.x = Buffer((n, IN), np.uint16)
        b.t = Buffer((n, OUT), np.float32)
        b.m = Buffer((n, OUT), np.float32)
        return b

    def take(self, idx: np.ndarray, out: Batch) -> Batch:
        """Rows idx of this batch into ``out`` (sized len(idx))."""
        ix = Buffer.of(np.asarray(idx, np.int64))
        check(api().w2d_mlp_take(self.x.ptr, self.t.ptr, self.m.ptr, ix.ptr, int(len(idx)), out.x.ptr, out.t.ptr,
                                 out.m.ptr, None))  # fmt: skip
        return out


class Net:
    def __init__(self, params: np.ndarray | None = None, seed: int = 0) -> None:
        self.nparam, self.nbf16 = sizes()
        p = he_init(seed) if params is None else np.asarray(params, np.float32)
        if p.size != self.nparam:
            raise ValueError(f"{p.size} parameters for a network of {self.nparam}")
        self.params = Buffer.of(p)
        self.wbf = Buffer((self.nbf16,), np.uint16)
        self.m = Buffer.of(np.zeros(self.nparam, np.float32))
        self.v = Buffer.of(np.zeros(self.nparam, np.float32))
        self.grad = Buffer((self.nparam + 1,), np.float32)
        self._partial: Buffer | None = None
        self.steps = 0
        check(api().w2d_mlp_pack(self.params.ptr, self.wbf.ptr, None))

    def forward(self, batch: Batch) -> np.ndarray:
        """y [n][2] for the batch's inputs."""
        y = Buffer((batch.n, OUT), np.float32)
        check(api().w2d_mlp_forward(self.wbf.ptr, self.params.ptr, batch.x.ptr, batch.n, y.ptr, None))
        return y.download()

    def gradient(self, batch: Batch, norm: float | None = None) -> float:
        """Loss sum(mask * (y - target)^2) / norm (norm: the sample count by default); the gradient stays in
        ``self.grad`` for ``adam`` (``gradient_host`` reads it)."""
        blocks = ctypes.c_int32(0)
        check(api().w2d_mlp_train_blocks(batch.n, ctypes.byref(blocks)))
        size = blocks.value * (self.nparam + 1)
        if self._partial is None or self._partial.shape[0] < size:
            self._partial = Buffer((size,), np.float32)
        check(api().w2d_mlp_gradient(self.wbf.ptr, self.params.ptr, batch.x.ptr, batch.t.ptr, batch.m.ptr, batch.n,
                                     float(norm or batch.n), self._partial.ptr, blocks.value, self.grad.ptr, None))
        return float(self.gradient_host()[self.nparam])

    def gradient_host(self) -> np.ndarray:
        return self.grad.download()

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.steps += 1
        check(api().w2d_mlp_adam(self.params.ptr, self.m.ptr, self.v.ptr, self.grad.ptr, self.steps, lr, beta1, beta2,
                                 eps, self.wbf.ptr, None))  # fmt: skip

    def host_params(self) -> np.ndarray:
        return self.params.download()
