"""The run engine around the step: forcing, boundaries, the water ledger, the run command (wadi2d/engine, run.py)."""
