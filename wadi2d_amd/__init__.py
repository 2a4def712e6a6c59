"""MI355X-native uniform-grid shallow-water step for Wadi2D: HIP kernels behind the reference's backend protocol."""

__version__ = "0.1.0"
