"""The command line of the run protocol (wadi2d/engine/cli.py:70-87,477-491): ``run JOB --out DIR [--backend]
[--resume] [--dry-run]`` and ``doctor``. ``python -m wadi2d_amd run job.json --out folder``; installed, ``wadi2d``.

``--backend`` takes the reference's names (auto, gpu, cpu) and this package's (hip, hip-exact, hip-f64); auto and
gpu run the HIP step, cpu is the reference's own Numba path and ends in exit 4 here.
"""

from __future__ import annotations

import argparse
import json
import platform
import sys

EXIT_OK = 0


def _run(args: argparse.Namespace) -> int:
    from wadi2d_amd.job import JobError, load_job
    from wadi2d_amd.run import EXIT_DONE, EXIT_INVALID, build_grid, run_job

    try:
        cfg = load_job(args.job)
        if args.dry_run:
            tables = build_grid(cfg)
            grid = tables.grid
            print(json.dumps({"cells": int(grid.mask.sum()), "tiles": int(grid.ntiles - 1), "dx": grid.dx,
                              "duration_s": cfg.run.duration_s}, indent=2))  # fmt: skip
            return EXIT_DONE
    except JobError as error:
        print(json.dumps({"ev": "error", "code": error.code, "message": str(error)}), file=sys.stderr)
        return error.code
    except ValueError as error:
        print(json.dumps({"ev": "error", "code": EXIT_INVALID, "message": str(error)}), file=sys.stderr)
        return EXIT_INVALID
    return run_job(cfg, args.out, args.backend, resume=args.resume, echo=True, loop=args.loop).exit_code


def _doctor(_args: argparse.Namespace) -> int:
    report: dict[str, object] = {"python": platform.python_version(), "platform": platform.platform()}
    try:
        from wadi2d_amd.solver.hip_backend import device_count

        report["hip_devices"] = device_count()
        report["hip_library"] = "loaded"
    except Exception as error:  # noqa: BLE001 - a report, never a failure
        report["hip_devices"] = 0
        report["hip_library"] = str(error)
    print(json.dumps(report, indent=2))
    return EXIT_OK


def parser() -> argparse.ArgumentParser:
    top = argparse.ArgumentParser(prog="wadi2d")
    sub = top.add_subparsers(dest="command", required=True)
    sub.add_parser("doctor", help="environment report").set_defaults(call=_doctor)
    run = sub.add_parser("run", help="one job")
    run.add_argument("job")
    run.add_argument("--out", required=True)
    run.add_argument("--backend", choices=("auto", "gpu", "cpu", "hip", "hip-exact", "hip-f64"), default=None)
    run.add_argument("--dry-run", action="store_true")
    run.add_argument("--resume", action="store_true")
    run.add_argument("--loop", choices=("host", "device"), default="host",
                     help="host: the host steps the model, as the reference does; device: the steps run on the device "
                          "in chunks between outputs, the same results (a HIP backend)")  # fmt: skip
    run.set_defaults(call=_run)
    return top


def main(argv: list[str] | None = None) -> int:
    args = parser().parse_args(argv)
    return args.call(args)


if __name__ == "__main__":
    sys.exit(main())
