// Seeded synthetic ground for the benchmark and test scenarios (host C++, OpenMP over rows).
//
// Value noise, fixed once here: for octave o = 0 .. n-1 the lattice spacing is base / 2^o and the weight p^o;
// lattice values are uniform in [-1, 1) from a splitmix64 hash of (seed, o, ix, iy); between lattice points the
// value is interpolated bilinearly with smoothstep weights 3t^2 - 2t^3; the octave sum is divided by the sum of
// the weights, so the noise lies in [-1, 1]. Ridged noise uses 1 - |v| per octave, rescaled to [-1, 1].
// Coordinates: pixel (r, c) of a block whose north-west corner is (x0, y0) has its centre at
// (x0 + (c + 0.5) p, y0 - (r + 0.5) p): rows run south, as the grid's do.

#include <cmath>
#include <cstdint>

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

inline double lattice(uint64_t seed, int oct, int64_t ix, int64_t iy) {
    uint64_t h = mix64(seed);
    h = mix64(h ^ (uint64_t)(oct + 1) * 0x100000001b3ULL);
    h = mix64(h ^ (uint64_t)ix * 0x9E3779B97F4A7C15ULL);
    h = mix64(h ^ (uint64_t)iy * 0xC2B2AE3D27D4EB4FULL);
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

inline double octave(uint64_t seed, int o, double x, double y, double wl) {
    const double fx = x / wl, fy = y / wl;
    const double gx = std::floor(fx), gy = std::floor(fy);
    const int64_t ix = (int64_t)gx, iy = (int64_t)gy;
    const double sx = smooth(fx - gx), sy = smooth(fy - gy);
    const double v00 = lattice(seed, o, ix, iy), v10 = lattice(seed, o, ix + 1, iy);
    const double v01 = lattice(seed, o, ix, iy + 1), v11 = lattice(seed, o, ix + 1, iy + 1);
    const double a = v00 + sx * (v10 - v00);
    const double b = v01 + sx * (v11 - v01);
    return a + sy * (b - a);
}

}  // namespace

extern "C" {

// out[r * nx + c] += amplitude * noise(pixel centre); ridged = 1 selects ridged noise.
void w2d_noise(double* out, int64_t ny, int64_t nx, double x0, double y0, double pixel, uint64_t seed, int32_t octaves,
               double base_wavelength, double persistence, double amplitude, int32_t ridged, int32_t threads) {
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (int64_t r = 0; r < ny; ++r) {
        const double y = y0 - (r + 0.5) * pixel;
        for (int64_t c = 0; c < nx; ++c) {
            const double x = x0 + (c + 0.5) * pixel;
            double sum = 0.0, norm = 0.0, w = 1.0, wl = base_wavelength;
            for (int o = 0; o < octaves; ++o) {
                double v = octave(seed, o, x, y, wl);
                if (ridged) v = 1.0 - 2.0 * std::fabs(v);
                sum += w * v;
                norm += w;
                w *= persistence;
                wl *= 0.5;
            }
            out[r * nx + c] += amplitude * (sum / norm);
        }
    }
}

}  // extern "C"
