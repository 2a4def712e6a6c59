"""Compile the package's native libraries in-tree (``wadi2d_amd/lib``).

``libwadi2d_host.so``: the table builder and the synthetic inputs, g++ with OpenMP.
``libwadi2d_hip.so``: the HIP step for gfx950 and its C-ABI, hipcc. Every kernel is built with packed FP32
instructions off and without FP contraction in the exact variants (see csrc/step.hip).
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "lib"
ROOT = HERE.parent
ARCH = os.environ.get("WADI2D_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["tables.cpp", "synth.cpp"]
HIP_SOURCES = ["step.hip", "mlp.hip", "tables_gpu.hip"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libwadi2d_host.so"
    srcs = [str(CSRC / s) for s in HOST_SOURCES if (CSRC / s).exists()]
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-o", str(out), *srcs])  # fmt: skip
    return out


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libwadi2d_hip.so"
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    _run(["/opt/rocm/bin/hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
          "-I", str(ROOT / "include"), "-I", str(CSRC), "-o", str(out), *srcs])  # fmt: skip
    return out


def build_all() -> None:
    build_host()
    build_hip()


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[which]()
