"""A run folder's files: progress.jsonl, the results and the checkpoint (wadi2d/io/progress.py, results.py,
checkpoint.py).

The reference writes ``results.h5`` and ``checkpoint.h5`` with h5py, which this image lacks; the same groups are kept
here as a folder of NumPy files with the reference's names, so a reader walks the same tree:

* ``results/attributes.json`` -- the root attributes (format, grid origin and cell size, the resolved job, and at
  the end the status, steps and water balance);
* ``results/mask.npy``; ``results/maxima/<name>.npy``; ``results/maps/t<seconds:09d>/<name>.npy`` with ``t_s`` in
  ``results/maps/t.../attributes.json``; ``results/series/<kind>/<name>.npz`` (``t_s`` and one array per quantity);
* ``checkpoint.npz`` -- state arrays, maxima, tally totals, series so far, the water accounts and the job's hash.

Maps are (ny, nx) float32 of the padded grid, rows north to south.
"""

from __future__ import annotations

import csv
import json
import re
import sys
from pathlib import Path
from typing import Any, TextIO

import numpy as np

FORMAT_VERSION = "1"
CHECKPOINT_FORMAT = "1"


class Progress:
    """Writes ``{"ev": name, ...}`` lines, flushed at once; ``echo`` copies them to stdout (io/progress.py:11-34)."""

    def __init__(self, path: str | Path, echo: bool = False, append: bool = False) -> None:
        self.path = Path(path)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self._file: TextIO = self.path.open("a" if append else "w", encoding="utf-8")
        self.echo = echo

    def emit(self, ev: str, **fields: Any) -> None:
        line = json.dumps({"ev": ev, **fields}, ensure_ascii=False, default=float)
        self._file.write(line + "\n")
        self._file.flush()
        if self.echo:
            print(line, file=sys.stdout, flush=True)

    def close(self) -> None:
        self._file.close()

    def __enter__(self) -> Progress:
        return self

    def __exit__(self, *_exc: object) -> None:
        self.close()


def safe_name(name: str) -> str:
    """A name fit for a file: letters, digits, dot, dash and underscore (io/results.py:35-37)."""
    return re.sub(r"[^A-Za-z0-9._-]+", "_", name).strip("_") or "unnamed"


class Results:
    """An open results folder; snapshots are written as they come, so a stopped run keeps what it wrote."""

    def __init__(self, folder: str | Path, grid, config_json: str, append: bool = False) -> None:
        self.root = Path(folder)
        self.root.mkdir(parents=True, exist_ok=True)
        attrs = self.root / "attributes.json"
        if append and attrs.exists():
            self.attrs = json.loads(attrs.read_text(encoding="utf-8"))
            return
        self.attrs = {"format_version": FORMAT_VERSION, "x0": grid.x0, "y0": grid.y0, "dx": grid.dx,
                      "config": config_json}  # fmt: skip
        np.save(self.root / "mask.npy", grid.mask)
        self._write_attrs()

    def _write_attrs(self) -> None:
        (self.root / "attributes.json").write_text(json.dumps(self.attrs, indent=1, default=float), encoding="utf-8")

    def snapshot(self, t: float, fields: dict[str, np.ndarray]) -> None:
        group = self.root / "maps" / f"t{round(t):09d}"
        group.mkdir(parents=True, exist_ok=True)
        for name, a in fields.items():
            np.save(group / f"{name}.npy", np.asarray(a, np.float32))
        (group / "attributes.json").write_text(json.dumps({"t_s": t}), encoding="utf-8")

    def maxima(self, fields: dict[str, np.ndarray]) -> None:
        group = self.root / "maxima"
        group.mkdir(parents=True, exist_ok=True)
        for name, a in fields.items():
            np.save(group / f"{name}.npy", np.asarray(a, np.float32))

    def series(self, kind: str, name: str, t: np.ndarray, columns: dict[str, np.ndarray]) -> None:
        group = self.root / "series" / kind
        group.mkdir(parents=True, exist_ok=True)
        arrays = {"t_s": np.asarray(t, np.float64), **{k: np.asarray(v, np.float64) for k, v in columns.items()}}
        np.savez(group / f"{safe_name(name)}.npz", **arrays)

    def attributes(self, values: dict[str, Any]) -> None:
        self.attrs.update(values)
        self._write_attrs()


def write_series_csv(path: str | Path, t, columns: dict[str, Any]) -> Path:
    """t_s plus one column per quantity (io/results.py:101-112)."""
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    names = list(columns)
    with out.open("w", newline="", encoding="utf-8") as f:
        writer = csv.writer(f)
        writer.writerow(["t_s", *names])
        for row in zip(t, *(columns[n] for n in names), strict=True):
            writer.writerow([f"{v:.6g}" if k else f"{v:.3f}" for k, v in enumerate(row)])
    return out


SAVED_FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om", "u1", "v1", "qx1", "qy1")
"""The state, the residuals, the wet share, and the second velocity and discharge buffers (io/checkpoint.py:17-19
saves the first ten; the second buffers make a resume bit for bit the unbroken run on tiles that sleep; the device
keeps one face momentum, so mx1 and my1 are mx and my)."""


def save_checkpoint(path: str | Path, payload: dict[str, Any]) -> Path:
    """Write through a temporary name; a crash during the write leaves the last good checkpoint
    (io/checkpoint.py:22-48)."""
    out = Path(path)
    tmp = out.with_name(out.stem + ".tmp.npz")
    arrays = {}
    meta = {"format": CHECKPOINT_FORMAT}
    for key, value in payload.items():
        if isinstance(value, np.ndarray):
            arrays[key] = value
        else:
            meta[key] = value
    arrays["meta"] = np.frombuffer(json.dumps(meta, default=float).encode(), np.uint8)
    np.savez(tmp, **arrays)
    tmp.replace(out)
    return out


def load_checkpoint(path: str | Path) -> dict[str, Any] | None:
    """The checkpoint's contents, or None for a file too old or damaged to read (io/checkpoint.py:51-60)."""
    try:
        with np.load(path, allow_pickle=False) as f:
            meta = json.loads(bytes(f["meta"]).decode())
            if meta.get("format") != CHECKPOINT_FORMAT:
                return None
            out: dict[str, Any] = dict(meta)
            for key in f.files:
                if key != "meta":
                    out[key] = f[key]
            return out
    except (OSError, KeyError, ValueError):
        return None
