"""Loading the package's native libraries. A missing library is an error, never a fallback."""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
HOST_LIB = LIB_DIR / "libwadi2d_host.so"
HIP_LIB = LIB_DIR / "libwadi2d_hip.so"

_loaded: dict[str, ctypes.CDLL] = {}


class NativeMissing(ImportError):
    """A native library of the package has not been built."""


def load(path: Path) -> ctypes.CDLL:
    key = str(path)
    if key not in _loaded:
        if not path.exists():
            raise NativeMissing(f"{path.name} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        _loaded[key] = ctypes.CDLL(key)
    return _loaded[key]


def host() -> ctypes.CDLL:
    return load(HOST_LIB)


def hip() -> ctypes.CDLL:
    return load(HIP_LIB)
