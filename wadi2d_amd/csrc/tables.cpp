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

#include "table_rows.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

using w2d_rows::CLOSED;
using w2d_rows::CRITICAL_WEIR;
using w2d_rows::TILE;

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
    hbar_at.resize(p.size());
    w2d_rows::cell_rows(p.data(), (int)p.size(), rows, tab, zmin, relmean, hc, hbar_at.data());
}

// core.py:44-79. prof is modified (crest). tab is (rows, 2). Returns invert zf and table top.
void face_table(std::vector<double>& prof, double crest, double pixel, double n_inv, double length, double e_lo,
                double e_hi, int rows, double* tab, double& zf, double& top) {
    w2d_rows::face_rows(prof.data(), (int)prof.size(), crest, pixel, n_inv, length, e_lo, e_hi, rows, tab, zf, top);
}

// core.py:82-179: overwrite the conveyance column with the local hydraulic radius.
struct RadiusScratch {
    std::vector<double> d, per, lo, hi, s0, s1, p0, p1;
    void size(int n) {
        d.assign(n, 0.0);
        per.assign(n, 0.0);
        lo.assign(n, 0.0);
        hi.assign(n, 0.0);
        s0.assign(n + 1, 0.0);
        s1.assign(n + 1, 0.0);
        p0.assign(n + 1, 0.0);
        p1.assign(n + 1, 0.0);
    }
};

void local_radius_conveyance(const double* ext, int n, int i0, int i1, double zf, double top, double pixel,
                             double n_inv, double length, double beta, int rows, double* tab, RadiusScratch& S) {
    S.size(n);
    w2d_rows::local_radius_rows(const_cast<double*>(ext), n, i0, i1, zf, top, pixel, n_inv, length, beta, rows, tab,
                                S.d.data(), S.per.data(), S.lo.data(), S.hi.data(), S.s0.data(), S.s1.data(),
                                S.p0.data(), S.p1.data());
}

double window_top(const double* line, int n, double zf, double top, double e_hi) {
    return w2d_rows::window_top(const_cast<double*>(line), n, zf, top, e_hi);
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
        std::vector<double> pix((size_t)k * k), hbar_at, prof(k), full(k + 2 * margin);
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
                    face_table(prof, xz, pixel, n_inv, crest_at(crest_xlen, gj, gi, 1.0), cfg->e_lo,
                               cfg->e_hi, rows, ftx.data(), zfx, efx);
                    // y face (north) (core.py:257-264)
                    for (int q = 0; q < k; ++q) prof[q] = std::max(z(r0 - 1, c0 + q), z(r0, c0 + q));
                    double zfy, efy;
                    const double yz = crest_at(crest_yz, gj, gi, NaN);
                    face_table(prof, yz, pixel, n_inv, crest_at(crest_ylen, gj, gi, 1.0), cfg->e_lo,
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
                                face_table(prof, NaN, pixel, n_inv, len, top, cfg->e_hi, rows,
                                           ftx.data(), zfx, efx);
                            }
                            const int e = std::min(margin, (int)std::ceil(cfg->beta * efx / pixel));
                            local_radius_conveyance(full.data() + margin - e, k + 2 * e, e, e + k, zfx, efx, pixel,
                                                    n_inv, len, cfg->beta, rows, ftx.data(), S);
                        }
                        if (std::isnan(yz)) {
                            const double len = crest_at(crest_ylen, gj, gi, 1.0);
                            for (int q = 0; q < nfull; ++q)
                                full[q] = std::max(zbig[(size_t)(margin + r0 - 1) * WB + c0 + q],
                                                   zbig[(size_t)(margin + r0) * WB + c0 + q]);
                            const double top = window_top(full.data(), nfull, zfy, efy, cfg->e_hi);
                            if (top > efy) {
                                for (int q = 0; q < k; ++q) prof[q] = full[margin + q];
                                face_table(prof, NaN, pixel, n_inv, len, top, cfg->e_hi, rows,
                                           fty.data(), zfy, efy);
                            }
                            const int e = std::min(margin, (int)std::ceil(cfg->beta * efy / pixel));
                            local_radius_conveyance(full.data() + margin - e, k + 2 * e, e, e + k, zfy, efy, pixel,
                                                    n_inv, len, cfg->beta, rows, fty.data(), S);
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
