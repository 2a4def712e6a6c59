"""The uniform-grid solver: backend protocol, host state, stepper and model (wadi2d/solver/)."""
