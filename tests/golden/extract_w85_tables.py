"""Copy the reference's golden sub-grid tables into this repository as test fixtures.

Source: ``W85 what limits the time step/builder_margin/tables_{long,short_flat,short_slope}.npz`` in the reference,
written by ``W85 what limits the time step/scripts/s08_builder_margin.py:29-42`` (three analytic beds on a 30 x 30
grid padded to one 32 x 32 tile, default ``Tables()``: 16 rows, local radius, Manning 0.03). These are data files
(inputs are closed-form beds, outputs the reference builder's arrays); they are loaded with numpy's pickle-free
loader and re-saved unchanged. Run once in a container that has the reference mounted:

    python tests/golden/extract_w85_tables.py /path/to/reference
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

SHA256 = {
    "tables_long.npz": "3c4f1b45d5332df9dadf2dbfb85c0423e379555e406de7671f9755dd77e93e97",
    "tables_short_flat.npz": "d12652114d11e88bc989e02eda4f56e968f35de53cf4f7a0180ca000c25350a9",
    "tables_short_slope.npz": "c4ec3a792488159a363ed48a6d04e9d501dd747f584b5c61b0ae221778bcd77b",
}


def main(reference: Path) -> None:
    src = reference / "W85 what limits the time step" / "builder_margin"
    here = Path(__file__).resolve().parent
    for name, digest in SHA256.items():
        raw = (src / name).read_bytes()
        if hashlib.sha256(raw).hexdigest() != digest:
            raise SystemExit(f"{name}: unexpected content")
        with np.load(src / name, allow_pickle=False) as f:
            arrays = {k: f[k] for k in f.files}
        np.savez_compressed(here / f"w85_{name}", **arrays)
        print(name, "->", here / f"w85_{name}")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
