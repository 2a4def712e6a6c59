// Sub-grid table builder: bed pixels -> every static array the step reads (host C++, OpenMP over tiles).
//
// Follows the reference's builder row for row, in float64, with the float32 cast at the very end:
//   cell table ............ wadi2d/tables/core.py:16-41   (_cell_table)
//   face table ............ wadi2d/tables/core.py:44-79   (_face_table)
//   local radius .......... wadi2d/tables/core.py:82-179  (_local_radius_conveyance, Spada et al. 2015 eqs. 14-17)
//   tile rows ............. wadi2d/tables/core.py:235-264, 324-351  (_tile_row, _local_radius_row)
//   window top ............ wadi2d/tables/core.py:354-371  (window_top, _raise_top)
//   centre-line beds ...... wadi2d/tables/core.py:267-279  (centre_beds)
//   driver, inverts ....... wadi2d/tables/builder.py:131-258, 261-269, 290-308  (build_tables, tile_margin, _faces)
// Each tile writes only its own slots, so the tables do not depend on the thread count (builder.py:147-148).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

constexpr int TILE = 32;
constexpr double EPS_RELIEF = 1e-9;          // core.py:13
constexpr double CLOSED = 1.0e9;             // solver/layout.py:41
constexpr double CRITICAL_WEIR = 1.7046;     // config/constants.py:6, (2/3) sqrt(2 g / 3)

struct Bed {  // builder.py:38-56 ArrayBed: reads clip to the edge pixel
    const double* z;
    int64_t ny, nx, row_off, col_off;
    double at(int64_t r, int64_t c) const {
        r = std::min(std::max(r + row_off, (int64_t)0), ny - 1);
        c = std::min(std::max(c + col_off, (int64_t)0), nx - 1);
        return z[r * nx + c];
    }
};

// core.py:16-41. Returns zmin, mean bed over zmin, Hc; fills tab[rows].
void cell_table(std::vector<double>& p, int rows, double* tab, double& zmin, double& relmean, double& hc,
                std::vector<double>& hbar_at) {
    std::sort(p.begin(), p.end());
    const int n = (int)p.size();
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
    hbar_at.resize(n);
    for (int m = 0; m < n; ++m) {
        cum += p[m];
        hbar_at[m] = (p[m] * (m + 1) - cum) / n;
    }
    const int last = rows - 1;
    int m = 0;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double hb = hc * (x * x);
        while (m + 1 < n && hbar_at[m + 1] <= hb) ++m;
        tab[r] = p[m] + (hb - hbar_at[m]) * n / (m + 1);
    }
    tab[last] = p[n - 1];
}

// core.py:44-79. prof is modified (crest). tab is (rows, 2). Returns invert zf and table top.
void face_table(std::vector<double>& prof, double crest, double pixel, double n_inv, const double* weight,
                double length, double e_lo, double e_hi, int rows, double* tab, double& zf, double& top) {
    const int m = (int)prof.size();
    if (!std::isnan(crest))
        for (int q = 0; q < m; ++q) prof[q] = std::max(prof[q], crest);
    zf = prof[0];
    for (int q = 1; q < m; ++q) zf = std::min(zf, prof[q]);
    double relmax = -std::numeric_limits<double>::infinity();
    for (int q = 0; q < m; ++q) relmax = std::max(relmax, prof[q] - zf);
    top = std::max(std::min(relmax, e_hi), e_lo);
    const int last = rows - 1;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double e = top * (x * x);
        double area = 0.0, conv = 0.0;
        for (int q = 0; q < m; ++q) {
            const double dep = e - (prof[q] - zf);
            if (dep > 0.0) {
                area += dep;
                const double c = std::cbrt(dep);
                conv += weight[q] * dep * (c * c);
            }
        }
        tab[2 * r] = pixel * area * length;
        tab[2 * r + 1] = std::pow(pixel * n_inv * conv * length, 0.6);
    }
}

// core.py:82-179: overwrite the conveyance column with the local hydraulic radius.
struct RadiusScratch {
    std::vector<double> d, per, s0, s1, p0, p1;
    std::vector<int64_t> run_lo, run_hi;
    void size(int n) {
        d.assign(n, 0.0);
        per.assign(n, 0.0);
        run_lo.assign(n, 0);
        run_hi.assign(n, 0);
        s0.assign(n + 1, 0.0);
        s1.assign(n + 1, 0.0);
        p0.assign(n + 1, 0.0);
        p1.assign(n + 1, 0.0);
    }
};

void local_radius_conveyance(const double* ext, int n, int i0, int i1, double zf, double top, double pixel,
                             double n_inv, const double* weight, double length, double beta, int rows, double* tab,
                             RadiusScratch& S) {
    S.size(n);
    const int last = rows - 1;
    const double pix2 = pixel * pixel;
    for (int r = 0; r < rows; ++r) {
        const double x = (double)r / last;
        const double level = zf + top * (x * x);
        for (int t = 0; t < n; ++t) {
            const double dep = level - ext[t];
            S.d[t] = dep > 0.0 ? dep : 0.0;
        }
        for (int t = 0; t < n; ++t) S.per[t] = 0.0;
        if (S.d[0] > 0.0) S.per[0] += 0.5 * pixel;
        if (S.d[n - 1] > 0.0) S.per[n - 1] += 0.5 * pixel;
        for (int t = 0; t < n - 1; ++t) {
            const bool wa = S.d[t] > 0.0, wb = S.d[t + 1] > 0.0;
            if (!(wa || wb)) continue;
            const double rise = ext[t + 1] - ext[t];
            const double seg = std::sqrt(pix2 + rise * rise);
            if (wa && wb) {
                S.per[t] += 0.5 * seg;
                S.per[t + 1] += 0.5 * seg;
            } else if (wa) {
                S.per[t] += rise > 0.0 ? std::min(S.d[t] / rise, 1.0) * seg : 0.5 * pixel;
            } else if (-rise > 0.0) {
                S.per[t + 1] += std::min(S.d[t + 1] / -rise, 1.0) * seg;
            } else {
                S.per[t + 1] += 0.5 * pixel;
            }
        }
        int64_t start = -1;
        for (int t = 0; t < n; ++t) {
            if (S.d[t] > 0.0) {
                if (start < 0) start = t;
                S.run_lo[t] = start;
            } else {
                start = -1;
            }
        }
        int64_t end = -1;
        for (int t = n - 1; t >= 0; --t) {
            if (S.d[t] > 0.0) {
                if (end < 0) end = t;
                S.run_hi[t] = end;
            } else {
                end = -1;
            }
        }
        for (int t = 0; t < n; ++t) {
            S.s0[t + 1] = S.s0[t] + S.d[t];
            S.s1[t + 1] = S.s1[t] + t * S.d[t];
            S.p0[t + 1] = S.p0[t] + S.per[t];
            S.p1[t + 1] = S.p1[t] + t * S.per[t];
        }
        double conv = 0.0;
        for (int i = i0; i < i1; ++i) {
            const double di = S.d[i];
            if (di <= 0.0) continue;
            const double w = std::max(beta * di / pixel, 1e-9);
            const int64_t lo = std::max(S.run_lo[i], (int64_t)std::ceil(i - w));
            const int64_t hi = std::min(S.run_hi[i], (int64_t)std::floor(i + w));
            double a_sum = (1.0 - i / w) * (S.s0[i + 1] - S.s0[lo]) + (S.s1[i + 1] - S.s1[lo]) / w;
            a_sum += (1.0 + i / w) * (S.s0[hi + 1] - S.s0[i + 1]) - (S.s1[hi + 1] - S.s1[i + 1]) / w;
            double p_sum = (1.0 - i / w) * (S.p0[i + 1] - S.p0[lo]) + (S.p1[i + 1] - S.p1[lo]) / w;
            p_sum += (1.0 + i / w) * (S.p0[hi + 1] - S.p0[i + 1]) - (S.p1[hi + 1] - S.p1[i + 1]) / w;
            const double radius = p_sum > 0.0 ? pixel * a_sum / p_sum : di;
            const double c = std::cbrt(radius);
            conv += weight[i - i0] * di * (c * c);
        }
        tab[2 * r + 1] = std::pow(pixel * n_inv * conv * length, 0.6);
    }
}

// core.py:354-362
double window_top(const double* line, int n, double zf, double top, double e_hi) {
    double hi = -std::numeric_limits<double>::infinity();
    for (int t = 0; t < n; ++t)
        if (line[t] > hi) hi = line[t];
    return std::max(top, std::min(hi - zf, e_hi));
}

struct Config {
    int32_t rows;
    int32_t local_radius;  // 1: Spada local radius (the default), 0: strips
    double e_lo, e_hi;     // face_depth_min, face_depth_max
    double beta;           // local_radius_beta
    double crest_tol;      // crest_tolerance
};

struct Out {
    double *zmin, *zbar;  // (ny, nx)
    float *hc, *dtab;
    float *flx, *frx, *glx, *grx, *efx, *cwx, *ftabx;
    float *fly, *fry, *gly, *gry, *efy, *cwy, *ftaby;
    float *zcx, *zcy;
};

struct Crest {  // builder.py:62-80 CrestFaces; NaN level means an untouched face
    const double *xz, *xlen, *xc, *yz, *ylen, *yc;
};

}  // namespace

extern "C" {

// build_tables (builder.py:131-258) for a uniform grid, no edits and no lining. Grid arrays are (ny, nx)
// row-major; tiles is (ntiles, 2) of (tj, ti) with entry 0 the void tile. manning is one value (nman_grid == 0)
// or an (ny, nx) grid. The crest arrays may be null (no terrain objects). Returns 0, or 1 on a bad argument.
int w2d_build_tables(const double* bed, int64_t bed_ny, int64_t bed_nx, int64_t row_off, int64_t col_off,
                     double pixel, int32_t ny, int32_t nx, double dx, const int32_t* tiles, int32_t ntiles,
                     const uint8_t* mask, const double* manning, int32_t nman_grid, const double* crest_xz,
                     const double* crest_xlen, const double* crest_xc, const double* crest_yz,
                     const double* crest_ylen, const double* crest_yc, const Config* cfg, int32_t threads,
                     Out* out) {
    const int k = (int)std::lround(dx / pixel);
    if (k < 1 || std::fabs(k * pixel - dx) > 1e-6 * dx) return 1;
    if (ny % TILE || nx % TILE || cfg->rows < 2) return 1;
    const int rows = cfg->rows;
    const Bed B{bed, bed_ny, bed_nx, row_off, col_off};
    // builder.py:261-269 tile_margin: beta * e_hi pixels, no further than the bed is long
    int margin = 0;
    if (cfg->local_radius) {
        margin = (int)std::ceil(cfg->beta * cfg->e_hi / pixel);
        margin = (int)std::min<int64_t>(margin, std::max(bed_ny, bed_nx));
    }
    const size_t ncell = (size_t)ny * nx;
    std::vector<double> zfx_g(ncell, 0.0), zfy_g(ncell, 0.0), rel_g(ncell, 0.0), zcx_g(ncell, 0.0),
        zcy_g(ncell, 0.0);
    for (size_t c = 0; c < ncell; ++c) out->zmin[c] = 0.0;
    auto nman_at = [&](int gj, int gi) { return nman_grid ? manning[(size_t)gj * nx + gi] : manning[0]; };
    auto crest_at = [&](const double* a, int gj, int gi, double dflt) {
        return a ? a[(size_t)gj * nx + gi] : dflt;
    };
    const double NaN = std::numeric_limits<double>::quiet_NaN();

    // tile 0 is the void: zeros everywhere, efx = efy = 1 (builder.py:171-176)
    const size_t T2 = (size_t)TILE * TILE;
    for (size_t c = 0; c < T2; ++c) {
        out->hc[c] = 0.0f;
        out->efx[c] = 1.0f;
        out->efy[c] = 1.0f;
    }
    std::fill(out->dtab, out->dtab + T2 * rows, 0.0f);
    std::fill(out->ftabx, out->ftabx + T2 * rows * 2, 0.0f);
    std::fill(out->ftaby, out->ftaby + T2 * rows * 2, 0.0f);

#pragma omp parallel num_threads(std::max(threads, 1))
    {
        const int W = TILE * k + 1;
        const int WB = W + 2 * margin;
        std::vector<double> zbig((size_t)WB * WB);
        std::vector<double> pix((size_t)k * k), hbar_at, prof(k), weight(k, 1.0), full(k + 2 * margin);
        std::vector<double> ctab(rows), ftx((size_t)rows * 2), fty((size_t)rows * 2);
        RadiusScratch S;
#pragma omp for schedule(dynamic, 1)
        for (int t = 1; t < ntiles; ++t) {
            const int tj = tiles[2 * t], ti = tiles[2 * t + 1];
            const int j0 = tj * TILE, i0 = ti * TILE;
            const int64_t r0b = (int64_t)j0 * k - 1 - margin, c0b = (int64_t)i0 * k - 1 - margin;
            for (int r = 0; r < WB; ++r)
                for (int c = 0; c < WB; ++c) zbig[(size_t)r * WB + c] = B.at(r0b + r, c0b + c);
            // z = zbig[margin : margin + W, margin : margin + W]
            auto z = [&](int r, int c) { return zbig[(size_t)(r + margin) * WB + (c + margin)]; };
            for (int j = 0; j < TILE; ++j) {
                for (int i = 0; i < TILE; ++i) {
                    const int gj = j0 + j, gi = i0 + i;
                    const size_t g = (size_t)gj * nx + gi;
                    const size_t s = (size_t)t * T2 + (size_t)j * TILE + i;
                    const int r0 = 1 + j * k, c0 = 1 + i * k;
                    // cell table (core.py:247-248)
                    for (int a = 0; a < k; ++a)
                        for (int b = 0; b < k; ++b) pix[(size_t)a * k + b] = z(r0 + a, c0 + b);
                    double zmin, relmean, hc;
                    cell_table(pix, rows, ctab.data(), zmin, relmean, hc, hbar_at);
                    out->zmin[g] = zmin;
                    rel_g[g] = relmean;
                    out->hc[s] = (float)hc;
                    for (int r = 0; r < rows; ++r) out->dtab[s * rows + r] = (float)ctab[r];
                    const double n_inv = 1.0 / nman_at(gj, gi);
                    // x face (west): the higher of the pixels either side (core.py:249-256)
                    for (int q = 0; q < k; ++q) prof[q] = std::max(z(r0 + q, c0 - 1), z(r0 + q, c0));
                    double zfx, efx;
                    const double xz = crest_at(crest_xz, gj, gi, NaN);
                    face_table(prof, xz, pixel, n_inv, weight.data(), crest_at(crest_xlen, gj, gi, 1.0), cfg->e_lo,
                               cfg->e_hi, rows, ftx.data(), zfx, efx);
                    // y face (north) (core.py:257-264)
                    for (int q = 0; q < k; ++q) prof[q] = std::max(z(r0 - 1, c0 + q), z(r0, c0 + q));
                    double zfy, efy;
                    const double yz = crest_at(crest_yz, gj, gi, NaN);
                    face_table(prof, yz, pixel, n_inv, weight.data(), crest_at(crest_ylen, gj, gi, 1.0), cfg->e_lo,
                               cfg->e_hi, rows, fty.data(), zfy, efy);
                    if (cfg->local_radius) {
                        // core.py:332-351; crest faces keep their table
                        const int nfull = k + 2 * margin;
                        if (std::isnan(xz)) {
                            const double len = crest_at(crest_xlen, gj, gi, 1.0);
                            for (int q = 0; q < nfull; ++q)
                                full[q] = std::max(zbig[(size_t)(r0 + q) * WB + margin + c0 - 1],
                                                   zbig[(size_t)(r0 + q) * WB + margin + c0]);
                            const double top = window_top(full.data(), nfull, zfx, efx, cfg->e_hi);
                            if (top > efx) {  // _raise_top, core.py:365-371
                                for (int q = 0; q < k; ++q) prof[q] = full[margin + q];
                                face_table(prof, NaN, pixel, n_inv, weight.data(), len, top, cfg->e_hi, rows,
                                           ftx.data(), zfx, efx);
                            }
                            const int e = std::min(margin, (int)std::ceil(cfg->beta * efx / pixel));
                            local_radius_conveyance(full.data() + margin - e, k + 2 * e, e, e + k, zfx, efx, pixel,
                                                    n_inv, weight.data(), len, cfg->beta, rows, ftx.data(), S);
                        }
                        if (std::isnan(yz)) {
                            const double len = crest_at(crest_ylen, gj, gi, 1.0);
                            for (int q = 0; q < nfull; ++q)
                                full[q] = std::max(zbig[(size_t)(margin + r0 - 1) * WB + c0 + q],
                                                   zbig[(size_t)(margin + r0) * WB + c0 + q]);
                            const double top = window_top(full.data(), nfull, zfy, efy, cfg->e_hi);
                            if (top > efy) {
                                for (int q = 0; q < k; ++q) prof[q] = full[margin + q];
                                face_table(prof, NaN, pixel, n_inv, weight.data(), len, top, cfg->e_hi, rows,
                                           fty.data(), zfy, efy);
                            }
                            const int e = std::min(margin, (int)std::ceil(cfg->beta * efy / pixel));
                            local_radius_conveyance(full.data() + margin - e, k + 2 * e, e, e + k, zfy, efy, pixel,
                                                    n_inv, weight.data(), len, cfg->beta, rows, fty.data(), S);
                        }
                    }
                    zfx_g[g] = zfx;
                    zfy_g[g] = zfy;
                    out->efx[s] = (float)efx;
                    out->efy[s] = (float)efy;
                    for (int r = 0; r < 2 * rows; ++r) {
                        out->ftabx[s * 2 * rows + r] = (float)ftx[r];
                        out->ftaby[s * 2 * rows + r] = (float)fty[r];
                    }
                    // centre_beds (core.py:267-279): lowest pixel of the middle pixel column / row
                    auto col_min = [&](int col) {
                        double v = z(r0, col);
                        for (int a = 1; a < k; ++a) v = std::min(v, z(r0 + a, col));
                        return v;
                    };
                    auto row_min = [&](int row) {
                        double v = z(row, c0);
                        for (int b = 1; b < k; ++b) v = std::min(v, z(row, c0 + b));
                        return v;
                    };
                    double cx, cy;
                    if (k % 2) {
                        cx = col_min(c0 + k / 2);
                        cy = row_min(r0 + k / 2);
                    } else {
                        cx = 0.5 * (col_min(c0 + k / 2 - 1) + col_min(c0 + k / 2));
                        cy = 0.5 * (row_min(r0 + k / 2 - 1) + row_min(r0 + k / 2));
                    }
                    zcx_g[g] = cx - zmin;
                    zcy_g[g] = cy - zmin;
                }
            }
        }
    }

    // builder.py:245-256
    for (size_t c = 0; c < ncell; ++c) out->zbar[c] = out->zmin[c] + rel_g[c];
    auto clean = [](double v) { return std::isfinite(v) ? std::max(v, 0.0) : 0.0; };
    // _faces (builder.py:290-308); np.roll wraps, as there
    auto faces = [&](int axis, const std::vector<double>& zf, const double* crest_c, float* fl, float* fr, float* gl,
                     float* gr, float* cw) {
        for (int t = 1; t < ntiles; ++t) {
            const int tj = tiles[2 * t], ti = tiles[2 * t + 1];
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int gj = tj * TILE + j, gi = ti * TILE + i;
                    const size_t g = (size_t)gj * nx + gi;
                    const int nj = axis == 0 ? (gj + ny - 1) % ny : gj;
                    const int ni = axis == 1 ? (gi + nx - 1) % nx : gi;
                    const size_t gn = (size_t)nj * nx + ni;
                    const bool open = mask[g] && mask[gn];
                    const double vfl = open ? zf[g] - out->zmin[gn] : CLOSED;
                    const double vfr = open ? zf[g] - out->zmin[g] : CLOSED;
                    const double vgl = open ? zf[g] - out->zbar[gn] : 0.0;
                    const double vgr = open ? zf[g] - out->zbar[g] : 0.0;
                    const bool natural = open && vgl > cfg->crest_tol && vgr > cfg->crest_tol;
                    const double cc = crest_c ? crest_c[g] : 0.0;
                    double vcw = cc != 0.0 ? cc : (natural ? CRITICAL_WEIR : 0.0);
                    if (!open) vcw = 0.0;
                    const size_t s = (size_t)t * T2 + (size_t)j * TILE + i;
                    fl[s] = (float)vfl;
                    fr[s] = (float)vfr;
                    gl[s] = (float)vgl;
                    gr[s] = (float)vgr;
                    cw[s] = (float)vcw;
                }
        }
        for (size_t c = 0; c < T2; ++c) fl[c] = fr[c] = gl[c] = gr[c] = cw[c] = 0.0f;
    };
    faces(1, zfx_g, crest_xc, out->flx, out->frx, out->glx, out->grx, out->cwx);
    faces(0, zfy_g, crest_yc, out->fly, out->fry, out->gly, out->gry, out->cwy);
    for (int t = 0; t < ntiles; ++t)
        for (int j = 0; j < TILE; ++j)
            for (int i = 0; i < TILE; ++i) {
                const size_t s = (size_t)t * T2 + (size_t)j * TILE + i;
                if (t == 0) {
                    out->zcx[s] = out->zcy[s] = 0.0f;
                    continue;
                }
                const size_t g = (size_t)(tiles[2 * t] * TILE + j) * nx + tiles[2 * t + 1] * TILE + i;
                out->zcx[s] = (float)clean(zcx_g[g]);
                out->zcy[s] = (float)clean(zcy_g[g]);
            }
    return 0;
}

}  // extern "C"
