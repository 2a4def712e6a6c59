"""The C-ABI boundary on the CPU: both libraries load and export every function their headers declare, and the
ctypes binding names no function the header does not declare (no compute calls: no GPU here)."""

from __future__ import annotations

import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^(?:int|void|const char\*)\s+(w2d_\w+)\s*\(", text, flags=re.M))


def test_the_step_library_exports_every_declared_function():
    from wadi2d_amd import _native

    names = _declared("wadi2d_hip.h")
    assert len(names) > 40
    lib = _native.hip()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_the_host_library_exports_every_declared_function():
    from wadi2d_amd import _native

    names = _declared("wadi2d_host.h")
    assert names == {"w2d_build_tables", "w2d_noise"}
    lib = _native.host()
    assert all(hasattr(lib, n) for n in names)


def test_the_binding_and_the_header_agree():
    from wadi2d_amd.solver import hip_backend

    names = _declared("wadi2d_hip.h")
    bound = set(hip_backend._SIGS) | {"w2d_last_error"}
    assert bound <= names, sorted(bound - names)
    unbound = names - bound
    assert not unbound, f"declared but not bound: {sorted(unbound)}"
