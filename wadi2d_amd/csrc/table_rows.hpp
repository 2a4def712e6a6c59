// The rows of one cell's and one face's sub-grid tables, in float64: one source for the host builder (tables.cpp,
// g++) and the device builder (tables_gpu.hip, hipcc), so the two follow the same statements.
//
//   cell table ............ wadi2d/tables/core.py:16-41   (_cell_table)
//   face table ............ wadi2d/tables/core.py:44-79   (_face_table)
//   local radius .......... wadi2d/tables/core.py:82-179  (_local_radius_conveyance, Spada et al. 2015 eqs. 14-17)
//   window top ............ wadi2d/tables/core.py:354-362 (window_top)
//
// Scratch arrays and the rows written are reached through accessors (a[q] -> double&): a plain pointer on the host,
// a view whose element q of thread i lies at base[q * stride + i] on the device (neighbouring threads read
// neighbouring words).
// min and max are std::min/std::max's comparisons, so NaN and signed zeros go as on the host.

#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

#ifdef __HIPCC__
#define W2D_HD __host__ __device__
#else
#define W2D_HD
#endif

namespace w2d_rows {

constexpr int TILE = 32;
constexpr double EPS_RELIEF = 1e-9;       // core.py:13
constexpr double CLOSED = 1.0e9;          // solver/layout.py:41
constexpr double CRITICAL_WEIR = 1.7046;  // config/constants.py:6, (2/3) sqrt(2 g / 3)

W2D_HD inline double dmin(double a, double b) { return b < a ? b : a; }  // std::min
W2D_HD inline double dmax(double a, double b) { return a < b ? b : a; }  // std::max

struct Strided {
    double* base;
    size_t stride;
    W2D_HD double& operator[](ptrdiff_t q) const { return base[(size_t)q * stride]; }
};

// core.py:16-41 on sorted pixels p[0..n): zmin, mean bed over zmin, Hc; tab[rows]; hb is scratch of n.
template <class A, class T>
W2D_HD void cell_rows(A p, int n, int rows, T tab, double& zmin, double& relmean, double& hc, A hb) {
    zmin = p[0];
    for (int q = 0; q < n; ++q) p[q] -= zmin;  // rel
    double sum = 0.0;
    for (int q = 0; q < n; ++q) sum += p[q];
    relmean = sum / n;
    hc = p[n - 1] - relmean;
    if (hc <= EPS_RELIEF) {
        for (int r = 0; r < rows; ++r) tab[r] = 0.0;
        hc = 0.0;
        return;
    }
    double cum = 0.0;
    for (int m = 0; m < n; ++m) {
        cum += p[m];
        hb[m] = (p[m] * (m + 1) - cum) / n;
    }
    const int last = rows - 1;
    int m = 0;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double h = hc * (x * x);
        while (m + 1 < n && hb[m + 1] <= h) ++m;
        tab[r] = p[m] + (h - hb[m]) * n / (m + 1);
    }
    tab[last] = p[n - 1];
}

// core.py:44-79 with unit weights. prof[0..m) is modified (crest). tab is (rows, 2). Returns invert zf and top.
template <class A, class T>
W2D_HD void face_rows(A prof, int m, double crest, double pixel, double n_inv, double length, double e_lo,
                      double e_hi, int rows, T tab, double& zf, double& top) {
    if (!std::isnan(crest))
        for (int q = 0; q < m; ++q) prof[q] = dmax(prof[q], crest);
    zf = prof[0];
    for (int q = 1; q < m; ++q) zf = dmin(zf, prof[q]);
    double relmax = -INFINITY;
    for (int q = 0; q < m; ++q) relmax = dmax(relmax, prof[q] - zf);
    top = dmax(dmin(relmax, e_hi), e_lo);
    const int last = rows - 1;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double e = top * (x * x);
        double area = 0.0, conv = 0.0;
        for (int q = 0; q < m; ++q) {
            const double dep = e - (prof[q] - zf);
            if (dep > 0.0) {
                area += dep;
                const double c = cbrt(dep);
                conv += dep * (c * c);
            }
        }
        tab[2 * r] = pixel * area * length;
        tab[2 * r + 1] = pow(pixel * n_inv * conv * length, 0.6);
    }
}

// core.py:82-179: the conveyance column of tab (rows, 2) from the local hydraulic radius over ext[0..n), the face's
// own pixels being [i0, i1). Scratch: d, per, run_lo, run_hi of n; s0, s1, p0, p1 of n + 1 (run bounds as doubles).
template <class A, class T>
W2D_HD void local_radius_rows(A ext, int n, int i0, int i1, double zf, double top, double pixel, double n_inv,
                              double length, double beta, int rows, T tab, A d, A per, A run_lo, A run_hi, A s0,
                              A s1, A p0, A p1) {
    const int last = rows - 1;
    const double pix2 = pixel * pixel;
    s0[0] = 0.0, s1[0] = 0.0, p0[0] = 0.0, p1[0] = 0.0;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double level = zf + top * (x * x);
        for (int t = 0; t < n; ++t) {
            const double dep = level - ext[t];
            d[t] = dep > 0.0 ? dep : 0.0;
        }
        for (int t = 0; t < n; ++t) per[t] = 0.0;
        if (d[0] > 0.0) per[0] += 0.5 * pixel;
        if (d[n - 1] > 0.0) per[n - 1] += 0.5 * pixel;
        for (int t = 0; t < n - 1; ++t) {
            const bool wa = d[t] > 0.0, wb = d[t + 1] > 0.0;
            if (!(wa || wb)) continue;
            const double rise = ext[t + 1] - ext[t];
            const double seg = sqrt(pix2 + rise * rise);
            if (wa && wb) {
                per[t] += 0.5 * seg;
                per[t + 1] += 0.5 * seg;
            } else if (wa) {
                per[t] += rise > 0.0 ? dmin(d[t] / rise, 1.0) * seg : 0.5 * pixel;
            } else if (-rise > 0.0) {
                per[t + 1] += dmin(d[t + 1] / -rise, 1.0) * seg;
            } else {
                per[t + 1] += 0.5 * pixel;
            }
        }
        int64_t start = -1;
        for (int t = 0; t < n; ++t) {
            if (d[t] > 0.0) {
                if (start < 0) start = t;
                run_lo[t] = (double)start;
            } else {
                start = -1;
            }
        }
        int64_t end = -1;
        for (int t = n - 1; t >= 0; --t) {
            if (d[t] > 0.0) {
                if (end < 0) end = t;
                run_hi[t] = (double)end;
            } else {
                end = -1;
            }
        }
        for (int t = 0; t < n; ++t) {
            s0[t + 1] = s0[t] + d[t];
            s1[t + 1] = s1[t] + t * d[t];
            p0[t + 1] = p0[t] + per[t];
            p1[t + 1] = p1[t] + t * per[t];
        }
        double conv = 0.0;
        for (int i = i0; i < i1; ++i) {
            const double di = d[i];
            if (di <= 0.0) continue;
            const double w = dmax(beta * di / pixel, 1e-9);
            const int64_t rl = (int64_t)run_lo[i], rh = (int64_t)run_hi[i];
            const int64_t wl = (int64_t)ceil(i - w), wh = (int64_t)floor(i + w);
            const int64_t lo = rl < wl ? wl : rl;
            const int64_t hi = wh < rh ? wh : rh;
            double a_sum = (1.0 - i / w) * (s0[i + 1] - s0[lo]) + (s1[i + 1] - s1[lo]) / w;
            a_sum += (1.0 + i / w) * (s0[hi + 1] - s0[i + 1]) - (s1[hi + 1] - s1[i + 1]) / w;
            double p_sum = (1.0 - i / w) * (p0[i + 1] - p0[lo]) + (p1[i + 1] - p1[lo]) / w;
            p_sum += (1.0 + i / w) * (p0[hi + 1] - p0[i + 1]) - (p1[hi + 1] - p1[i + 1]) / w;
            const double radius = p_sum > 0.0 ? pixel * a_sum / p_sum : di;
            const double c = cbrt(radius);
            conv += di * (c * c);
        }
        tab[2 * r + 1] = pow(pixel * n_inv * conv * length, 0.6);
    }
}

// core.py:354-362
template <class A>
W2D_HD double window_top(A line, int n, double zf, double top, double e_hi) {
    double hi = -INFINITY;
    for (int t = 0; t < n; ++t)
        if (line[t] > hi) hi = line[t];
    return dmax(top, dmin(hi - zf, e_hi));
}

}  // namespace w2d_rows
