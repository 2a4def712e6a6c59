// The sub-grid table builder on the device: build_tables (wadi2d/tables/builder.py:131-258) for a uniform grid, the
// same statements as the host builder (tables.cpp) through the shared rows of table_rows.hpp, in float64, one thread
// per own cell.
//
//  * table_cells_kernel: a cell's table (its k x k pixels sorted by insertion), its west (x) and north (y) face tables
//    with the top window and the local-radius conveyance, and its centre-line beds. Pixels are read from the bed in
//    HBM (reads clip to its edge as the host's); every thread's scratch (pixels, profiles, the radius sums, its rows)
//    lies in one HBM buffer with thread i's element q at q * stride + i, so a wave's scratch accesses are coalesced.
//    Tiles go in batches that keep the scratch under 1 GiB.
//  * table_faces_kernel: the faces' inverts over each side's zmin and zbar and their natural crests
//    (builder.py:290-308), from the first kernel's grids.
//
// The device's cbrt, pow and sqrt are not the host library's, so a row may differ from the host builder's in its
// last bit before the float32 cast; the tests hold the two builders to one float32 ulp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "table_rows.hpp"
#include "wadi2d_hip.h"
#include "wadi2d_host.h"

namespace {

using namespace w2d_rows;
constexpr int T2 = TILE * TILE;
constexpr int NT = 256;

struct BedView {  // builder.py:38-56 ArrayBed: reads clip to the edge pixel
    const double* z;
    int64_t ny, nx, row_off, col_off;
    __device__ double at(int64_t r, int64_t c) const {
        r = r + row_off;
        c = c + col_off;
        r = r < 0 ? 0 : (r > ny - 1 ? ny - 1 : r);
        c = c < 0 ? 0 : (c > nx - 1 ? nx - 1 : c);
        return z[r * nx + c];
    }
};

struct CellArgs {
    BedView B;
    double pixel, e_lo, e_hi, beta;
    int k, margin, rows, nx, local_radius, nman_grid;
    const int32_t* tiles;
    int t_first, t_count;
    const double* manning;
    const double *cxz, *cxlen, *cyz, *cylen;  // crest levels and lengths, or null
    double* scratch;
    size_t stride;  // threads of the batch
    int per_thread;
    // outputs: grids (ny * nx) and tile arrays
    double *zmin, *rel, *zfx, *zfy, *zcx, *zcy;
    float *hc, *dtab, *efx, *efy, *ftabx, *ftaby;
};

__global__ __launch_bounds__(NT) void table_cells_kernel(CellArgs A) {
    const size_t idx = blockIdx.x * (size_t)NT + threadIdx.x;
    if (idx >= (size_t)A.t_count * T2) return;
    const int t = A.t_first + (int)(idx / T2);
    const int p = (int)(idx % T2), j = p / TILE, i = p - (p / TILE) * TILE;
    const int tj = A.tiles[2 * t], ti = A.tiles[2 * t + 1];
    const int64_t j0 = (int64_t)tj * TILE, i0 = (int64_t)ti * TILE;
    const int gj = (int)j0 + j, gi = (int)i0 + i;
    const size_t g = (size_t)gj * A.nx + gi;
    const size_t s = (size_t)t * T2 + p;
    const int k = A.k, m = A.margin, rows = A.rows;
    const int r0 = 1 + j * k, c0 = 1 + i * k;
    // the host's z(r, c) over its staged block, and zb(R, C) over the block with the margin
    auto z = [&](int r, int c) { return A.B.at(j0 * k - 1 + r, i0 * k - 1 + c); };
    auto zb = [&](int r, int c) { return A.B.at(j0 * k - 1 - m + r, i0 * k - 1 - m + c); };
    size_t off = 0;
    auto view = [&](int len) {
        const Strided v{A.scratch + off * A.stride + (idx), A.stride};
        off += (size_t)len;
        return v;
    };
    const int nfull = k + 2 * m;
    const Strided pix = view(k * k), hb = view(k * k), prof = view(k), full = view(nfull);
    const Strided d = view(nfull), per = view(nfull), lo = view(nfull), hi = view(nfull);
    const Strided s0 = view(nfull + 1), s1 = view(nfull + 1), p0 = view(nfull + 1), p1 = view(nfull + 1);
    const Strided ctab = view(rows), ftx = view(2 * rows), fty = view(2 * rows);

    // cell table (core.py:247-248): the cell's pixels, sorted
    const int np = k * k;
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) pix[a * k + b] = z(r0 + a, c0 + b);
    for (int q = 1; q < np; ++q) {
        const double v = pix[q];
        int w = q - 1;
        while (w >= 0 && v < pix[w]) {
            pix[w + 1] = pix[w];
            --w;
        }
        pix[w + 1] = v;
    }
    double zmin, relmean, hcv;
    cell_rows(pix, np, rows, ctab, zmin, relmean, hcv, hb);
    A.zmin[g] = zmin;
    A.rel[g] = relmean;
    A.hc[s] = (float)hcv;
    for (int r = 0; r < rows; ++r) A.dtab[s * rows + r] = (float)ctab[r];
    const double n_inv = 1.0 / (A.nman_grid ? A.manning[g] : A.manning[0]);
    const double NaN = nan("");
    const double xz = A.cxz ? A.cxz[g] : NaN, yz = A.cyz ? A.cyz[g] : NaN;
    const double xlen = A.cxlen ? A.cxlen[g] : 1.0, ylen = A.cylen ? A.cylen[g] : 1.0;
    // x face (west): the higher of the pixels either side (core.py:249-256)
    for (int q = 0; q < k; ++q) prof[q] = dmax(z(r0 + q, c0 - 1), z(r0 + q, c0));
    double zfx, efx;
    face_rows(prof, k, xz, A.pixel, n_inv, xlen, A.e_lo, A.e_hi, rows, ftx, zfx, efx);
    // y face (north) (core.py:257-264)
    for (int q = 0; q < k; ++q) prof[q] = dmax(z(r0 - 1, c0 + q), z(r0, c0 + q));
    double zfy, efy;
    face_rows(prof, k, yz, A.pixel, n_inv, ylen, A.e_lo, A.e_hi, rows, fty, zfy, efy);
    if (A.local_radius) {
        // core.py:332-351; crest faces keep their table
        if (std::isnan(xz)) {
            for (int q = 0; q < nfull; ++q) full[q] = dmax(zb(r0 + q, m + c0 - 1), zb(r0 + q, m + c0));
            const double top = window_top(full, nfull, zfx, efx, A.e_hi);
            if (top > efx) {  // _raise_top, core.py:365-371
                for (int q = 0; q < k; ++q) prof[q] = full[m + q];
                face_rows(prof, k, NaN, A.pixel, n_inv, xlen, top, A.e_hi, rows, ftx, zfx, efx);
            }
            const int e = (int)dmin((double)m, ceil(A.beta * efx / A.pixel));
            const Strided ext{full.base + (size_t)(m - e) * A.stride, A.stride};
            local_radius_rows(ext, k + 2 * e, e, e + k, zfx, efx, A.pixel, n_inv, xlen, A.beta, rows, ftx, d, per,
                              lo, hi, s0, s1, p0, p1);
        }
        if (std::isnan(yz)) {
            for (int q = 0; q < nfull; ++q) full[q] = dmax(zb(m + r0 - 1, c0 + q), zb(m + r0, c0 + q));
            const double top = window_top(full, nfull, zfy, efy, A.e_hi);
            if (top > efy) {
                for (int q = 0; q < k; ++q) prof[q] = full[m + q];
                face_rows(prof, k, NaN, A.pixel, n_inv, ylen, top, A.e_hi, rows, fty, zfy, efy);
            }
            const int e = (int)dmin((double)m, ceil(A.beta * efy / A.pixel));
            const Strided ext{full.base + (size_t)(m - e) * A.stride, A.stride};
            local_radius_rows(ext, k + 2 * e, e, e + k, zfy, efy, A.pixel, n_inv, ylen, A.beta, rows, fty, d, per,
                              lo, hi, s0, s1, p0, p1);
        }
    }
    A.zfx[g] = zfx;
    A.zfy[g] = zfy;
    A.efx[s] = (float)efx;
    A.efy[s] = (float)efy;
    for (int r = 0; r < 2 * rows; ++r) {
        A.ftabx[s * 2 * rows + r] = (float)ftx[r];
        A.ftaby[s * 2 * rows + r] = (float)fty[r];
    }
    // centre_beds (core.py:267-279): lowest pixel of the middle pixel column / row
    auto col_min = [&](int col) {
        double v = z(r0, col);
        for (int a = 1; a < k; ++a) v = dmin(v, z(r0 + a, col));
        return v;
    };
    auto row_min = [&](int row) {
        double v = z(row, c0);
        for (int b = 1; b < k; ++b) v = dmin(v, z(row, c0 + b));
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
    A.zcx[g] = cx - zmin;
    A.zcy[g] = cy - zmin;
}

struct FaceArgs {
    int ny, nx, ntiles;
    const int32_t* tiles;
    const uint8_t* mask;
    double crest_tol;
    const double *zmin, *rel, *zfx, *zfy, *zcx, *zcy, *cxc, *cyc;
    double* zbar;
    float *flx, *frx, *glx, *grx, *cwx, *fly, *fry, *gly, *gry, *cwy, *zcxo, *zcyo;
};

__device__ inline double clean(double v) { return isfinite(v) ? dmax(v, 0.0) : 0.0; }

// builder.py:245-256, 290-308: zbar; each own face's inverts over the two sides' zmin and zbar and its crest
// coefficient; np.roll wraps, as there.
__global__ __launch_bounds__(NT) void table_faces_kernel(FaceArgs F) {
    const size_t idx = blockIdx.x * (size_t)NT + threadIdx.x;
    if (idx >= (size_t)F.ntiles * T2) return;
    const int t = (int)(idx / T2), p = (int)(idx % T2), j = p / TILE, i = p - (p / TILE) * TILE;
    const size_t s = idx;
    if (t == 0) {
        F.flx[s] = F.frx[s] = F.glx[s] = F.grx[s] = F.cwx[s] = 0.0f;
        F.fly[s] = F.fry[s] = F.gly[s] = F.gry[s] = F.cwy[s] = 0.0f;
        F.zcxo[s] = F.zcyo[s] = 0.0f;
        return;
    }
    const int gj = F.tiles[2 * t] * TILE + j, gi = F.tiles[2 * t + 1] * TILE + i;
    const size_t g = (size_t)gj * F.nx + gi;
    auto zbar = [&](size_t c) { return F.zmin[c] + F.rel[c]; };
    for (int axis = 1; axis >= 0; --axis) {
        const int nj = axis == 0 ? (gj + F.ny - 1) % F.ny : gj;
        const int ni = axis == 1 ? (gi + F.nx - 1) % F.nx : gi;
        const size_t gn = (size_t)nj * F.nx + ni;
        const double zf = axis ? F.zfx[g] : F.zfy[g];
        const bool open = F.mask[g] && F.mask[gn];
        const double vfl = open ? zf - F.zmin[gn] : CLOSED;
        const double vfr = open ? zf - F.zmin[g] : CLOSED;
        const double vgl = open ? zf - zbar(gn) : 0.0;
        const double vgr = open ? zf - zbar(g) : 0.0;
        const bool natural = open && vgl > F.crest_tol && vgr > F.crest_tol;
        const double* crest_c = axis ? F.cxc : F.cyc;
        const double cc = crest_c ? crest_c[g] : 0.0;
        double vcw = cc != 0.0 ? cc : (natural ? CRITICAL_WEIR : 0.0);
        if (!open) vcw = 0.0;
        if (axis) {
            F.flx[s] = (float)vfl, F.frx[s] = (float)vfr, F.glx[s] = (float)vgl, F.grx[s] = (float)vgr;
            F.cwx[s] = (float)vcw;
        } else {
            F.fly[s] = (float)vfl, F.fry[s] = (float)vfr, F.gly[s] = (float)vgl, F.gry[s] = (float)vgr;
            F.cwy[s] = (float)vcw;
        }
    }
    F.zcxo[s] = (float)clean(F.zcx[g]);
    F.zcyo[s] = (float)clean(F.zcy[g]);
}

__global__ void zbar_kernel(size_t n, const double* __restrict__ zmin, const double* __restrict__ rel,
                            double* __restrict__ zbar) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        zbar[c] = zmin[c] + rel[c];
}

struct Pool {  // device buffers of one build, zeroed, freed on every exit; the first failure is kept
    std::vector<void*> held;
    hipError_t err = hipSuccess;
    ~Pool() {
        for (void* p : held) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (err != hipSuccess) return nullptr;
        const size_t bytes = (count ? count : 1) * sizeof(T);
        if ((err = hipMalloc(&p, bytes)) != hipSuccess) return nullptr;
        held.push_back(p);
        err = hipMemset(p, 0, bytes);
        return (T*)p;
    }
};

}  // namespace

int w2d_report(int code, const char* msg);  // csrc/step.hip: the message w2d_last_error returns

extern "C" {

int w2d_build_tables_device(const double* bed, int64_t bed_ny, int64_t bed_nx, int64_t row_off, int64_t col_off,
                            double pixel, int32_t ny, int32_t nx, double dx, const int32_t* tiles, int32_t ntiles,
                            const uint8_t* mask, const double* manning, int32_t nman_grid, const double* crest_xz,
                            const double* crest_xlen, const double* crest_xc, const double* crest_yz,
                            const double* crest_ylen, const double* crest_yc, const w2d_table_config* cfg,
                            w2d_table_out* out) {
    if (!bed || !tiles || !mask || !manning || !cfg || !out || bed_ny < 1 || bed_nx < 1 || ntiles < 1)
        return w2d_report(1, "build_tables_device: bad arguments");
    const int k = (int)std::lround(dx / pixel);
    if (k < 1 || std::fabs(k * pixel - dx) > 1e-6 * dx || k > 64)
        return w2d_report(1, "build_tables_device: the cell is not 1 to 64 whole pixels");
    if (ny % TILE || nx % TILE || cfg->rows < 2 || cfg->rows > 64)
        return w2d_report(1, "build_tables_device: bad grid or row count");
    for (int t = 1; t < ntiles; ++t)
        if (tiles[2 * t] < 0 || tiles[2 * t + 1] < 0 || (tiles[2 * t] + 1) * TILE > ny || (tiles[2 * t + 1] + 1) * TILE > nx)
            return w2d_report(1, "build_tables_device: a tile lies outside the grid");
    const int rows = cfg->rows;
    int margin = 0;  // builder.py:261-269 tile_margin
    if (cfg->local_radius) {
        margin = (int)std::ceil(cfg->beta * cfg->e_hi / pixel);
        margin = (int)std::min<int64_t>(margin, std::max(bed_ny, bed_nx));
    }
    const size_t ncell = (size_t)ny * nx, nt = (size_t)ntiles;
    hipError_t e = hipSuccess;
    Pool P;
    double* d_bed = P.get<double>((size_t)bed_ny * bed_nx);
    int32_t* d_tiles = P.get<int32_t>(2 * nt);
    uint8_t* d_mask = P.get<uint8_t>(ncell);
    double* d_man = P.get<double>(nman_grid ? ncell : 1);
    const double* crest_in[6] = {crest_xz, crest_xlen, crest_xc, crest_yz, crest_ylen, crest_yc};
    double* d_crest[6] = {};
    for (int q = 0; q < 6; ++q)
        if (crest_in[q]) d_crest[q] = P.get<double>(ncell);
    double *g_zmin = P.get<double>(ncell), *g_rel = P.get<double>(ncell), *g_zbar = P.get<double>(ncell);
    double *g_zfx = P.get<double>(ncell), *g_zfy = P.get<double>(ncell);
    double *g_zcx = P.get<double>(ncell), *g_zcy = P.get<double>(ncell);
    float *t_hc = P.get<float>(nt * T2), *t_dtab = P.get<float>(nt * T2 * rows);
    float *t_efx = P.get<float>(nt * T2), *t_efy = P.get<float>(nt * T2);
    float *t_ftabx = P.get<float>(nt * T2 * rows * 2), *t_ftaby = P.get<float>(nt * T2 * rows * 2);
    float* t_face[12];
    for (int q = 0; q < 12; ++q) t_face[q] = P.get<float>(nt * T2);
    if (P.err != hipSuccess)
        return w2d_report(P.err == hipErrorOutOfMemory ? 2 : 3, hipGetErrorString(P.err));
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    if ((e = up(d_bed, bed, (size_t)bed_ny * bed_nx * 8)) != hipSuccess ||
        (e = up(d_tiles, tiles, 2 * nt * 4)) != hipSuccess || (e = up(d_mask, mask, ncell)) != hipSuccess ||
        (e = up(d_man, manning, (nman_grid ? ncell : 1) * 8)) != hipSuccess)
        return w2d_report(3, hipGetErrorString(e));
    for (int q = 0; q < 6; ++q)
        if (crest_in[q] && (e = up(d_crest[q], crest_in[q], ncell * 8)) != hipSuccess)
            return w2d_report(3, hipGetErrorString(e));
    // scratch: per thread 2 k^2 + k + 9 (k + 2 m) + 4 + 5 rows doubles; batches of tiles keep it under 1 GiB
    const int nfull = k + 2 * margin;
    const int per = 2 * k * k + k + nfull + 4 * nfull + 4 * (nfull + 1) + 5 * rows;
    const size_t budget = (size_t)1 << 30;
    const int batch = (int)std::max<size_t>(1, std::min<size_t>(nt, budget / ((size_t)T2 * per * 8)));
    double* scratch = P.get<double>((size_t)batch * T2 * per);
    if (P.err != hipSuccess)
        return w2d_report(P.err == hipErrorOutOfMemory ? 2 : 3, hipGetErrorString(P.err));
    CellArgs A{};
    A.B = BedView{d_bed, bed_ny, bed_nx, row_off, col_off};
    A.pixel = pixel, A.e_lo = cfg->e_lo, A.e_hi = cfg->e_hi, A.beta = cfg->beta;
    A.k = k, A.margin = margin, A.rows = rows, A.nx = nx, A.local_radius = cfg->local_radius;
    A.nman_grid = nman_grid, A.tiles = d_tiles, A.manning = d_man;
    A.cxz = d_crest[0], A.cxlen = d_crest[1], A.cyz = d_crest[3], A.cylen = d_crest[4];
    A.scratch = scratch, A.per_thread = per;
    A.zmin = g_zmin, A.rel = g_rel, A.zfx = g_zfx, A.zfy = g_zfy, A.zcx = g_zcx, A.zcy = g_zcy;
    A.hc = t_hc, A.dtab = t_dtab, A.efx = t_efx, A.efy = t_efy, A.ftabx = t_ftabx, A.ftaby = t_ftaby;
    for (int t0 = 1; t0 < ntiles; t0 += batch) {
        A.t_first = t0;
        A.t_count = std::min(batch, ntiles - t0);
        A.stride = (size_t)A.t_count * T2;
        const unsigned blocks = (unsigned)((A.stride + NT - 1) / NT);
        table_cells_kernel<<<blocks, NT>>>(A);
        if ((e = hipGetLastError()) != hipSuccess) return w2d_report(3, hipGetErrorString(e));
    }
    zbar_kernel<<<1024, NT>>>(ncell, g_zmin, g_rel, g_zbar);
    FaceArgs F{ny, nx, ntiles, d_tiles, d_mask, cfg->crest_tol, g_zmin, g_rel, g_zfx, g_zfy, g_zcx, g_zcy,
               d_crest[2], d_crest[5], g_zbar, t_face[0], t_face[1], t_face[2], t_face[3], t_face[4], t_face[5],
               t_face[6], t_face[7], t_face[8], t_face[9], t_face[10], t_face[11]};
    table_faces_kernel<<<(unsigned)((nt * T2 + NT - 1) / NT), NT>>>(F);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess)
        return w2d_report(3, hipGetErrorString(e));
    // tile 0 is the void: zeros, efx = efy = 1 (builder.py:171-176)
    std::vector<float> ones(T2, 1.0f);
    if ((e = up(t_efx, ones.data(), T2 * 4)) != hipSuccess || (e = up(t_efy, ones.data(), T2 * 4)) != hipSuccess)
        return w2d_report(3, hipGetErrorString(e));
    auto down = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    };
    float* outs[12] = {out->flx, out->frx, out->glx, out->grx, out->cwx, out->fly,
                       out->fry, out->gly, out->gry, out->cwy, out->zcx, out->zcy};
    if ((e = down(out->zmin, g_zmin, ncell * 8)) != hipSuccess || (e = down(out->zbar, g_zbar, ncell * 8)) != hipSuccess ||
        (e = down(out->hc, t_hc, nt * T2 * 4)) != hipSuccess ||
        (e = down(out->dtab, t_dtab, nt * T2 * rows * 4)) != hipSuccess ||
        (e = down(out->efx, t_efx, nt * T2 * 4)) != hipSuccess || (e = down(out->efy, t_efy, nt * T2 * 4)) != hipSuccess ||
        (e = down(out->ftabx, t_ftabx, nt * T2 * rows * 8)) != hipSuccess ||
        (e = down(out->ftaby, t_ftaby, nt * T2 * rows * 8)) != hipSuccess)
        return w2d_report(3, hipGetErrorString(e));
    for (int q = 0; q < 12; ++q)
        if ((e = down(outs[q], t_face[q], nt * T2 * 4)) != hipSuccess) return w2d_report(3, hipGetErrorString(e));
    return 0;
}

}  // extern "C"
