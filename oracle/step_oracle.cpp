// CPU restatement of Wadi2D's uniform-grid time step. TEST INFRASTRUCTURE ONLY: the checker of the HIP path
// and the timed CPU baseline, never part of the product.
//
// It follows the reference's Numba twin statement for statement, in the reference's own array layout
// (32 x 32 own cells inside a 2-cell halo, STRIDE 36, void tile 0, halos filled by an exchange map):
//   face step ............ wadi2d/solver/cpu/kernels.py:86-219   (faces.cu:105-242)
//   faces ................ wadi2d/solver/cpu/kernels.py:222-244
//   limit ................ wadi2d/solver/cpu/kernels.py:247-262
//   update ............... wadi2d/solver/cpu/kernels.py:265-299
//   plan ................. wadi2d/solver/cpu/kernels.py:302-311
//   exchange ............. wadi2d/solver/cpu/kernels.py:314-319, map wadi2d/solver/state.py:14-48
//   boundary ............. wadi2d/solver/cpu/kernels.py:322-334
//   step order ........... wadi2d/solver/cpu/backend.py:95-126
//   helpers .............. wadi2d/solver/cpu/kernels.py:19-83 (minmod, ustar, gated, cell_level, cell_plan,
//                          face_lookup)
// One template over the arithmetic type: double is the FP64 oracle, float (same statement order, same
// float32 arithmetic as the Numba twin) is the CPU baseline. OpenMP runs tiles in parallel as Numba's prange
// does; every tile writes only its own entries, so the result does not depend on the thread count.
//
// Parity pins: see tests/test_oracle_*.py (closed forms, invariants, the reference's golden tables).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr int TILE = 32;
constexpr int HALO = 2;
constexpr int STRIDE = TILE + 2 * HALO;
constexpr int STRIDE2 = STRIDE * STRIDE;
constexpr int TILE2 = TILE * TILE;

// wadi2d/grid/tiles.py:12 NEIGHBOURS, in the order of TileGrid.nbr
constexpr int NB_DJ[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
constexpr int NB_DI[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

struct Params {
    double g, dx, e_wet, h_cfl, u_max, lead, a_min;
    int rows, second, sheet, mirror, carry;
};

// np.cbrt on the type's own values; for float the double result rounded once to float.
inline double cbrt_of(double x) { return std::cbrt(x); }
inline float cbrt_of(float x) { return (float)std::cbrt((double)x); }

// int(s) of a non-negative s below `cap`, as the reference's min(int(s), last - 1); NaN gives row 0 (the
// reference leaves it undefined, a NaN state stops the run at the next refresh either way).
template <class R>
inline int row_of(R s, int cap) {
    if (!(s > R(0))) return 0;
    return s < R(cap) ? (int)s : cap;
}

template <class R>
inline R minmod(R a, R b) {  // kernels.py:20-23
    if (a * b <= R(0)) return R(0);
    return std::fabs(a) < std::fabs(b) ? a : b;
}

template <class R>
inline R ustar(R uc, R uup, R udown, bool second) {  // kernels.py:27-30
    if (!second) return uc;
    return uc + R(0.5) * minmod(uc - uup, udown - uc);
}

template <class R>
inline R gated(R neighbour, R own, bool mirror) {  // kernels.py:34-36
    return (mirror && neighbour == R(0)) ? own : neighbour;
}

template <class R>
inline R cell_level(R hb, R hc, const float* tab, int rows) {  // kernels.py:40-48
    if (hc <= R(0)) return hb;
    const int last = rows - 1;
    if (hb >= hc) return R(tab[last]) + (hb - hc);
    const R s = R(std::sqrt(hb / hc)) * R(last);
    const int r = row_of(s, last - 1);
    return R(tab[r]) + (s - R(r)) * (R(tab[r + 1]) - R(tab[r]));
}

template <class R>
inline R cell_plan(R hb, R hc, const float* tab, int rows) {  // kernels.py:52-59
    if (hc <= R(0) || hb >= hc) return R(1);
    const int last = rows - 1;
    const R s = R(std::sqrt(hb / hc)) * R(last);
    const int r = row_of(s, last - 1);
    const R m = R(last);
    return hc * R(2 * r + 1) / (m * m) / std::max(R(tab[r + 1]) - R(tab[r]), R(1e-12));
}

template <class R>
inline void face_lookup(R e, R top, const float* tab, int rows, R& area, R& t, R& w) {  // kernels.py:63-83
    const int last = rows - 1;
    const R m = R(last);
    if (e >= top) {
        const R a1 = R(tab[2 * last]);
        const R a0 = R(tab[2 * last - 2]);
        const R c = (m - R(1)) / m;
        w = (a1 - a0) / (top * (R(1) - c * c));
        area = a1 + w * (e - top);
        t = R(tab[2 * last + 1]) * (area / std::max(a1, R(1e-12)));
        return;
    }
    const R s = R(std::sqrt(e / top)) * m;
    const int r = row_of(s, last - 1);
    const R c0 = R(r) / m;
    const R c1 = R(r + 1) / m;
    const R e0 = c0 * c0 * top;
    const R de = (c1 * c1 - c0 * c0) * top;
    const R f = (e - e0) / de;
    const R a0 = R(tab[2 * r]);
    const R a1 = R(tab[2 * r + 2]);
    const R t0 = R(tab[2 * r + 1]);
    area = a0 + f * (a1 - a0);
    t = t0 + f * (R(tab[2 * r + 3]) - t0);
    w = (a1 - a0) / de;
}

template <class R>
struct FaceStatics {
    const float *fl, *fr, *gl, *gr, *ef, *cw, *ftab;
};

// kernels.py:86-219. s: array step along the normal, p across it. Returns the wave speed.
template <class R>
R face_step(int k, int fid, R dt, bool want_wave, const R* om, const R* zc, int s, int p, const R* h, const R* d,
            const R* u, const R* m, const R* q, const R* qp, const R* vp, const FaceStatics<R>& st, R* u1, R* m1,
            R* q1, const Params& P) {
    const int rows = P.rows;
    const R g = R(P.g), dx = R(P.dx), e_wet = R(P.e_wet), h_cfl = R(P.h_cfl), u_max = R(P.u_max);
    const R lead = R(P.lead), a_min = R(P.a_min);
    const bool second = P.second, sheet = P.sheet, mirror = P.mirror, carry = P.carry;
    const int far = 2;
    const R span = dx;
    const R e_l = d[k - s] - R(st.fl[fid]);
    const R e_r = d[k] - R(st.fr[fid]);
    const R uf = u[k];
    const R e_lp = std::max(e_l, R(0));
    const R e_rp = std::max(e_r, R(0));
    const bool from_left = uf > R(0) || (uf == R(0) && e_l >= e_r);
    const R eup = from_left ? e_lp : e_rp;
    if (eup <= e_wet) {
        u1[k] = R(0);
        m1[k] = R(0);
        q1[k] = R(0);
        return R(0);
    }
    R eface = eup;
    if (carry && R(st.cw[fid]) == R(0)) {
        const R dz = from_left ? (zc[k - s] - R(st.fl[fid])) : (zc[k] - R(st.fr[fid]));
        if (dz > R(1e-6)) {
            R phi = (eup - R(2) * dz) / (R(2) * dz);
            phi = phi < R(0) ? R(0) : (phi > R(1) ? R(1) : phi);
            eface = eup - phi * dz;
        }
    }
    const float* tab = st.ftab + (size_t)2 * rows * fid;
    R area, t, w;
    face_lookup(eface, R(st.ef[fid]), tab, rows, area, t, w);
    const R h_l = h[k - s];
    const R h_r = h[k];
    const R hup = from_left ? h_l : h_r;

    // [1] section
    R rho = R(1);
    R d_eta = e_rp - e_lp;
    if (sheet) {
        const R avol = hup * dx;
        const R edown = from_left ? e_rp : e_lp;
        if (avol < area) {
            const R sub = std::min(e_lp, e_rp) / std::max(e_lp, e_rp);
            const R sub2 = sub * sub;
            const R alpha = sub2 * sub2;
            rho = alpha + (R(1) - alpha) * avol / area;
        }
        if (edown <= R(0)) {
            // [2] disconnected
            const R g_u = from_left ? R(st.gl[fid]) : R(st.gr[fid]);
            const R g_d = from_left ? R(st.gr[fid]) : R(st.gl[fid]);
            const R h_d = from_left ? h_r : h_l;
            const R drop = (hup - g_u) - std::max(h_d - g_d, R(0.5) * (g_u - g_d));
            const R level_drop = std::fabs(e_l - e_r);
            const R drive = std::max(eup, std::min(drop, level_drop));
            d_eta = from_left ? -drive : drive;
        }
    }

    // [3] momentum fluxes
    const R um = gated(u[k - s], uf, mirror);
    const R up = gated(u[k + s], uf, mirror);
    const R umm = gated(u[k - 2 * s], um, mirror);
    const R upp = gated(u[k + far * s], up, mirror);
    const R qc_l = R(0.5) * (q[k - s] + q[k]);
    const R qc_r = R(0.5) * (q[k] + q[k + s]);
    const R s_l = qc_l > R(0) ? ustar(um, umm, uf, second) : ustar(uf, up, um, second);
    const R s_r = qc_r > R(0) ? ustar(uf, um, up, second) : ustar(up, upp, uf, second);
    R flux = qc_r * s_r - qc_l * s_l;
    const R qy_s = R(0.5) * (qp[k - s] + qp[k]);
    const R qy_n = R(0.5) * (qp[k - s + p] + qp[k + p]);
    const R u_s = gated(u[k - p], uf, mirror);
    const R u_n = gated(u[k + p], uf, mirror);
    const R uss = gated(u[k - 2 * p], u_s, mirror);
    const R unn = gated(u[k + 2 * p], u_n, mirror);
    const R us_s = qy_s > R(0) ? ustar(u_s, uss, uf, second) : ustar(uf, u_n, u_s, second);
    const R us_n = qy_n > R(0) ? ustar(uf, u_s, u_n, second) : ustar(u_n, unn, uf, second);
    flux += qy_n * us_n - qy_s * us_s;

    // [4] momentum and pressure
    const R abar = R(0.5) * (h_l + h_r) * dx;
    const R press = dt * g * d_eta / span;
    R us;
    if (abar > R(1e-6) * dx)
        us = (m[k] - dt * flux / span) / abar - press;
    else
        us = uf - press;

    // [5] friction
    const R vb = R(0.25) * (vp[k - s] + vp[k] + vp[k - s + p] + vp[k + p]);
    const R t3 = t * t * t * cbrt_of(t);
    R fric = area * area / std::max(t3, R(1e-30));
    if (rho < R(1)) fric = fric / std::max(rho * cbrt_of(rho), R(1e-12));
    const R a = dt * g * fric;
    const R sp = std::sqrt(us * us + vb * vb);
    R un = R(2) * us / (R(1) + std::sqrt(R(1) + R(4) * a * sp));
    if (std::fabs(un) > u_max) un = un > R(0) ? u_max : -u_max;

    // [6] crest face
    R qn = area * rho * un;
    R c = R(st.cw[fid]);
    if (c != R(0)) {
        R area_c = area, w_c = w;
        if (c < R(0)) {
            c = -c;
            const R ua = from_left ? um : up;
            R t_c;
            face_lookup(eup + R(0.5) * ua * ua / g, R(st.ef[fid]), tab, rows, area_c, t_c, w_c);
        }
        const R qc = c * area_c * std::sqrt(area_c / std::max(w_c, R(1e-9)));
        if (std::fabs(qn) > qc) {
            qn = qn > R(0) ? qc : -qc;
            un = qn / std::max(area * rho, R(1e-12));
        }
    }
    u1[k] = un;
    m1[k] = abar * un;
    q1[k] = qn;
    if (!want_wave) return R(0);
    if (!std::isfinite(un)) return R(1e30);
    const R d_left = d[k - s];
    const R d_right = d[k];
    const R dup = from_left ? d_left : d_right;
    if (dup <= h_cfl) return R(0);
    R plan = from_left ? om[k - s] : om[k];
    const R dother = from_left ? d_right : d_left;
    if (dother > h_cfl) plan = std::min(plan, from_left ? om[k] : om[k - s]);
    const R wave = std::sqrt(std::max(g * area * rho / (std::max(plan, a_min) * dx), R(0)));
    return std::fabs(un) + std::min(lead * std::fabs(un - uf), R(2) * wave) + wave;
}

template <class R>
R scaled(R q, R rl, R rr) {  // kernels.py:266-267
    return q * (q > R(0) ? rl : rr);
}

enum Field {
    F_H, F_D, F_U, F_V, F_MX, F_MY, F_QX, F_QY, F_U1, F_V1, F_MX1, F_MY1, F_QX1, F_QY1, F_R, F_SRC, F_RES, F_OM,
    F_COUNT
};

template <class R>
struct Model {
    using Real = R;
    Params P;
    int ntiles = 0;
    int nhalo = 0;
    std::vector<R> a[F_COUNT];
    std::vector<R> zcx, zcy;  // state layout, halos filled once (state.py static_cell_state)
    std::vector<float> hc, dtab, flx, frx, glx, grx, efx, cwx, ftabx, fly, fry, gly, gry, efy, cwy, ftaby;
    std::vector<int> slot, src;  // exchange map: slot[nhalo] offsets in a tile, src[(ntiles-1)*nhalo]
    std::vector<R> wave;
    std::vector<int32_t> wet;
    std::vector<int32_t> active;
    std::vector<int32_t> bcell, bref, bkind;
    std::vector<float> bd0, bh0;
    int threads = 1;

    void build_exchange(const int32_t* nbr) {  // state.py:14-48
        std::vector<int> hj, hi;
        for (int jj = 0; jj < STRIDE; ++jj)
            for (int ii = 0; ii < STRIDE; ++ii) {
                const bool own = jj >= HALO && jj < HALO + TILE && ii >= HALO && ii < HALO + TILE;
                if (!own) {
                    hj.push_back(jj);
                    hi.push_back(ii);
                }
            }
        nhalo = (int)hj.size();
        slot.resize(nhalo);
        src.assign((size_t)std::max(ntiles - 1, 0) * nhalo, 0);
        for (int n = 0; n < nhalo; ++n) slot[n] = hj[n] * STRIDE + hi[n];
        for (int t = 1; t < ntiles; ++t)
            for (int n = 0; n < nhalo; ++n) {
                const int dj = hj[n] < HALO ? -1 : (hj[n] >= HALO + TILE ? 1 : 0);
                const int di = hi[n] < HALO ? -1 : (hi[n] >= HALO + TILE ? 1 : 0);
                int dir = 0;
                for (int q = 0; q < 8; ++q)
                    if (NB_DJ[q] == dj && NB_DI[q] == di) dir = q;
                const int sj = hj[n] - dj * TILE, si = hi[n] - di * TILE;
                src[(size_t)(t - 1) * nhalo + n] = nbr[(size_t)t * 8 + dir] * STRIDE2 + sj * STRIDE + si;
            }
    }

    void exchange(R* arr) { exchange_list(active.data(), (int)active.size(), arr); }

    void exchange_list(const int32_t* tiles, int n, R* arr) {  // kernels.py:314-319; own entries into halo entries
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int a = 0; a < n; ++a) {
            const int t = tiles[a];
            for (int m = 0; m < nhalo; ++m) arr[(size_t)t * STRIDE2 + slot[m]] = arr[src[(size_t)(t - 1) * nhalo + m]];
        }
    }

    void faces(R dt, bool want_wave) { faces_list(active.data(), (int)active.size(), dt, want_wave); }

    void faces_list(const int32_t* tiles, int n, R dt, bool want_wave) {  // kernels.py:222-244
        const FaceStatics<R> sx{flx.data(), frx.data(), glx.data(), grx.data(), efx.data(), cwx.data(), ftabx.data()};
        const FaceStatics<R> sy{fly.data(), fry.data(), gly.data(), gry.data(), efy.data(), cwy.data(), ftaby.data()};
        R *h = a[F_H].data(), *d = a[F_D].data(), *u = a[F_U].data(), *v = a[F_V].data();
        R *mx = a[F_MX].data(), *my = a[F_MY].data(), *qx = a[F_QX].data(), *qy = a[F_QY].data();
        R *u1 = a[F_U1].data(), *v1 = a[F_V1].data(), *mx1 = a[F_MX1].data(), *my1 = a[F_MY1].data();
        R *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data(), *om = a[F_OM].data();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (int ai = 0; ai < n; ++ai) {
            const int t = tiles[ai];
            R w = R(0);
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = t * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    const int fid = t * TILE2 + j * TILE + i;
                    const R wx = face_step<R>(k, fid, dt, want_wave, om, zcx.data(), 1, STRIDE, h, d, u, mx, qx, qy, v,
                                              sx, u1, mx1, qx1, P);
                    const R wy = face_step<R>(k, fid, dt, want_wave, om, zcy.data(), STRIDE, 1, h, d, v, my, qy, qx, u,
                                              sy, v1, my1, qy1, P);
                    w = std::max(w, std::max(wx, wy));
                }
            if (want_wave) wave[t] = w;
        }
    }

    void limit(R dt) { limit_list(active.data(), (int)active.size(), dt); }

    void limit_list(const int32_t* tiles, int n, R dt) {  // kernels.py:247-262
        const R dx = R(P.dx);
        const R area = dx * dx;
        const R *h = a[F_H].data(), *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data();
        R* r = a[F_R].data();
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int ai = 0; ai < n; ++ai) {
            const int t = tiles[ai];
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = t * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    const R out = std::max(qx1[k + 1], R(0)) + std::max(-qx1[k], R(0)) + std::max(qy1[k + STRIDE], R(0)) +
                                  std::max(-qy1[k], R(0));
                    const R have = std::max(h[k], R(0)) * area;
                    r[k] = out * dt > have ? have / (out * dt) : R(1);
                }
        }
    }

    void update(R dt, bool want_plan) {  // debt off in the uniform step (cpu/backend.py:432-435)
        update_list(active.data(), (int)active.size(), dt, want_plan, false);
    }

    // kernels.py:270-299; `debt` keeps an overdrawn cell's shortfall in its residual (the level stepper's form)
    void update_list(const int32_t* tiles, int n, R dt, bool want_plan, bool debt) {
        const R dx = R(P.dx);
        const R area = dx * dx;
        const int rows = P.rows;
        R *res = a[F_RES].data(), *om = a[F_OM].data(), *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data();
        R *r = a[F_R].data(), *srcv = a[F_SRC].data(), *h = a[F_H].data(), *d = a[F_D].data();
        R *qx = a[F_QX].data(), *qy = a[F_QY].data();
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int ai = 0; ai < n; ++ai) {
            const int t = tiles[ai];
            int any_wet = 0;
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = t * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    const int fid = t * TILE2 + j * TILE + i;
                    const R rc = r[k];
                    const R qw = scaled(qx1[k], r[k - 1], rc);
                    const R qe = scaled(qx1[k + 1], rc, r[k + 1]);
                    const R qn = scaled(qy1[k], r[k - STRIDE], rc);
                    const R qs = scaled(qy1[k + STRIDE], rc, r[k + STRIDE]);
                    qx[k] = qw;
                    qy[k] = qn;
                    const R h0 = h[k];
                    const R inc = res[k] + dt * srcv[k] - dt * (qe - qw + qs - qn) / area;
                    const R total = h0 + inc;
                    const R hn = std::max(total, R(0));
                    res[k] = total > R(0) ? inc - (total - h0) : (debt ? total : R(0));
                    h[k] = hn;
                    const float* tab = dtab.data() + (size_t)rows * fid;
                    d[k] = cell_level(hn, R(hc[fid]), tab, rows);
                    if (want_plan) om[k] = cell_plan(hn, R(hc[fid]), tab, rows);
                    if (hn > R(0)) any_wet = 1;
                }
            wet[t] = any_wet;
        }
    }

    void plan(const int32_t* tiles, int n) {  // kernels.py:302-311
        const int rows = P.rows;
        R *om = a[F_OM].data(), *h = a[F_H].data();
        for (int ai = 0; ai < n; ++ai) {
            const int t = tiles[ai];
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = t * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    const int fid = t * TILE2 + j * TILE + i;
                    om[k] = cell_plan(h[k], R(hc[fid]), dtab.data() + (size_t)rows * fid, rows);
                }
        }
    }

    void boundary() {  // kernels.py:322-334
        R *h = a[F_H].data(), *d = a[F_D].data();
        for (size_t m = 0; m < bcell.size(); ++m) {
            const int k = bcell[m];
            if (bkind[m] == 0) {
                h[k] = R(bh0[m]);
                d[k] = R(bd0[m]);
            } else if (bkind[m] == 1) {
                h[k] = R(0);
                d[k] = R(0);
            } else {
                h[k] = h[bref[m]];
                d[k] = std::max(d[bref[m]] + R(bd0[m]), R(0));
            }
        }
    }

    void step(double dt64, bool want_wave, bool want_wet) {  // cpu/backend.py:95-126
        (void)want_wet;  // the CPU twin writes every active tile's wet flag on every step
        if (want_wave) std::fill(wave.begin(), wave.end(), R(0));
        if (active.empty()) return;
        const R dt = R(dt64);
        faces(dt, want_wave);
        exchange(a[F_U1].data());
        exchange(a[F_V1].data());
        exchange(a[F_QX1].data());
        exchange(a[F_QY1].data());
        limit(dt);
        exchange(a[F_R].data());
        update(dt, want_wave);
        if (!bcell.empty()) boundary();
        exchange(a[F_H].data());
        exchange(a[F_D].data());
        exchange(a[F_QX].data());
        exchange(a[F_QY].data());
        if (want_wave) exchange(a[F_OM].data());
        std::swap(a[F_U], a[F_U1]);
        std::swap(a[F_V], a[F_V1]);
        std::swap(a[F_MX], a[F_MX1]);
        std::swap(a[F_MY], a[F_MY1]);
    }

    // ---- local time stepping: the level stepper's operations (wadi2d/solver/cpu/levels.py:1-147), over tile lists
    // and edge rows (fine tile, side of the fine tile 0 W 1 N 2 E 3 S, coarse tile, edge number)
    std::vector<R> h0, d0, est;
    std::vector<int32_t> level;

    void lts_init() {
        if (!h0.empty()) return;
        h0.assign((size_t)ntiles * STRIDE2, R(0));
        d0.assign((size_t)ntiles * STRIDE2, R(0));
        est.assign((size_t)4 * ntiles * TILE, R(0));
        level.assign(ntiles, 0);
    }

    static int edge_face(int t, int side, int q) {  // cpu/levels.py:12-21
        const int base = t * STRIDE2;
        if (side == 0) return base + (q + HALO) * STRIDE + HALO;
        if (side == 1) return base + HALO * STRIDE + q + HALO;
        if (side == 2) return base + (q + HALO) * STRIDE + HALO + TILE;
        return base + (HALO + TILE) * STRIDE + q + HALO;
    }

    static int edge_cell(int c, int side, int q) {  // cpu/levels.py:24-33
        const int base = c * STRIDE2;
        if (side == 0) return base + (q + HALO) * STRIDE + HALO + TILE - 1;
        if (side == 1) return base + (HALO + TILE - 1) * STRIDE + q + HALO;
        if (side == 2) return base + (q + HALO) * STRIDE + HALO;
        return base + HALO * STRIDE + q + HALO;
    }

    void strips(const int32_t* e, int ne, R dt, bool want_wave) {  // cpu/levels.py:36-57
        const FaceStatics<R> sx{flx.data(), frx.data(), glx.data(), grx.data(), efx.data(), cwx.data(), ftabx.data()};
        const FaceStatics<R> sy{fly.data(), fry.data(), gly.data(), gry.data(), efy.data(), cwy.data(), ftaby.data()};
        R *h = a[F_H].data(), *d = a[F_D].data(), *u = a[F_U].data(), *v = a[F_V].data();
        R *mx = a[F_MX].data(), *my = a[F_MY].data(), *qx = a[F_QX].data(), *qy = a[F_QY].data();
        R *u1 = a[F_U1].data(), *v1 = a[F_V1].data(), *mx1 = a[F_MX1].data(), *my1 = a[F_MY1].data();
        R *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data(), *om = a[F_OM].data();
        for (int n = 0; n < ne; ++n) {
            const int fine = e[4 * n], side = e[4 * n + 1], t = e[4 * n + 2];
            for (int q = 0; q < TILE; ++q) {
                R w;
                if (side == 2) {
                    const int k = t * STRIDE2 + (q + HALO) * STRIDE + HALO;
                    w = face_step<R>(k, t * TILE2 + q * TILE, dt, want_wave, om, zcx.data(), 1, STRIDE, h, d, u, mx, qx,
                                     qy, v, sx, u1, mx1, qx1, P);
                } else {
                    const int k = t * STRIDE2 + HALO * STRIDE + q + HALO;
                    w = face_step<R>(k, t * TILE2 + q, dt, want_wave, om, zcy.data(), STRIDE, 1, h, d, v, my, qy, qx,
                                     u, sy, v1, my1, qy1, P);
                }
                if (want_wave) {
                    wave[t] = std::max(wave[t], w);
                    wave[fine] = std::max(wave[fine], w);
                }
            }
        }
    }

    // cpu/levels.py:60-80: the halos of the listed tiles, a source tile in the middle of its step read at time n
    // between its old and new values
    void halo_levels(const int32_t* tiles, int nt, int n) {
        R *h = a[F_H].data(), *d = a[F_D].data(), *u = a[F_U].data(), *v = a[F_V].data();
        R *qx = a[F_QX].data(), *qy = a[F_QY].data();
        for (int x = 0; x < nt; ++x) {
            const int t = tiles[x];
            for (int m = 0; m < nhalo; ++m) {
                const int dst = t * STRIDE2 + slot[m];
                const int s = src[(size_t)(t - 1) * nhalo + m];
                const int span = 1 << level[s / STRIDE2];
                const int part = n & (span - 1);
                if (part == 0) {
                    h[dst] = h[s];
                    d[dst] = d[s];
                } else {
                    const R w = R(part) / R(span);
                    h[dst] = h0[s] + w * (h[s] - h0[s]);
                    d[dst] = d0[s] + w * (d[s] - d0[s]);
                }
                u[dst] = u[s];
                v[dst] = v[s];
                qx[dst] = qx[s];
                qy[dst] = qy[s];
            }
        }
    }

    void save_old(const int32_t* tiles, int nt) {  // cpu/levels.py:83-90
        const R *h = a[F_H].data(), *d = a[F_D].data();
        for (int x = 0; x < nt; ++x)
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = tiles[x] * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    h0[k] = h[k];
                    d0[k] = d[k];
                }
    }

    void commit(const int32_t* tiles, int nt) {  // cpu/levels.py:93-102
        for (int x = 0; x < nt; ++x)
            for (int j = 0; j < TILE; ++j)
                for (int i = 0; i < TILE; ++i) {
                    const int k = tiles[x] * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                    a[F_U][k] = a[F_U1][k];
                    a[F_V][k] = a[F_V1][k];
                    a[F_MX][k] = a[F_MX1][k];
                    a[F_MY][k] = a[F_MY1][k];
                }
    }

    void commit_strips(const int32_t* e, int ne) {  // cpu/levels.py:105-116
        for (int n = 0; n < ne; ++n) {
            const int t = e[4 * n + 2];
            for (int q = 0; q < TILE; ++q) {
                if (e[4 * n + 1] == 2) {
                    const int k = t * STRIDE2 + (q + HALO) * STRIDE + HALO;
                    a[F_U][k] = a[F_U1][k];
                    a[F_MX][k] = a[F_MX1][k];
                } else {
                    const int k = t * STRIDE2 + HALO * STRIDE + q + HALO;
                    a[F_V][k] = a[F_V1][k];
                    a[F_MY][k] = a[F_MY1][k];
                }
            }
        }
    }

    void edge_estimate(const int32_t* e, int ne) {  // cpu/levels.py:119-127
        const R *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data(), *r = a[F_R].data();
        for (int n = 0; n < ne; ++n) {
            const int side = e[4 * n + 1];
            for (int q = 0; q < TILE; ++q) {
                const int k = edge_face(e[4 * n], side, q);
                const int s = (side & 1) ? STRIDE : 1;
                const R f = (side & 1) ? qy1[k] : qx1[k];
                est[(size_t)e[4 * n + 3] * TILE + q] = scaled(f, r[k - s], r[k]);
            }
        }
    }

    void edge_flux(const int32_t* e, int ne, R dt) {  // cpu/levels.py:130-147
        const R dx = R(P.dx);
        const R *qx1 = a[F_QX1].data(), *qy1 = a[F_QY1].data(), *r = a[F_R].data();
        R *res = a[F_RES].data(), *qx = a[F_QX].data(), *qy = a[F_QY].data();
        for (int n = 0; n < ne; ++n) {
            const int side = e[4 * n + 1], c = e[4 * n + 2];
            for (int q = 0; q < TILE; ++q) {
                const int k = edge_face(e[4 * n], side, q);
                const int s = (side & 1) ? STRIDE : 1;
                const R f = (side & 1) ? qy1[k] : qx1[k];
                const R qs = scaled(f, r[k - s], r[k]);
                const R sign = side < 2 ? R(-1) : R(1);
                res[edge_cell(c, side, q)] += sign * dt * (qs - est[(size_t)e[4 * n + 3] * TILE + q]) / (dx * dx);
                if (side == 2) qx[c * STRIDE2 + (q + HALO) * STRIDE + HALO] = qs;
                if (side == 3) qy[c * STRIDE2 + HALO * STRIDE + q + HALO] = qs;
            }
        }
    }

    void refresh_halos() {  // cpu/backend.py:411-414
        if (active.empty()) return;
        for (Field f : {F_H, F_D, F_QX, F_QY, F_U, F_V, F_OM}) exchange(a[f].data());
    }
};

struct Tables {
    int32_t ntiles, rows;
    const int32_t* nbr;  // (ntiles, 8)
    const float *hc, *dtab;
    const float *flx, *frx, *glx, *grx, *efx, *cwx, *ftabx;
    const float *fly, *fry, *gly, *gry, *efy, *cwy, *ftaby;
    const float *zcx, *zcy;  // (ntiles, TILE, TILE)
};

template <class R>
Model<R>* create(const Tables* T, const Params* P, int threads) {
    auto* m = new Model<R>();
    m->P = *P;
    m->P.rows = T->rows;
    m->threads = std::max(threads, 1);
    m->ntiles = T->ntiles;
    const size_t n = (size_t)T->ntiles * STRIDE2;
    const size_t nc = (size_t)T->ntiles * TILE2;
    for (auto& v : m->a) v.assign(n, R(0));
    auto cp = [&](std::vector<float>& dst, const float* s, size_t count) { dst.assign(s, s + count); };
    cp(m->hc, T->hc, nc);
    cp(m->dtab, T->dtab, nc * T->rows);
    cp(m->flx, T->flx, nc); cp(m->frx, T->frx, nc); cp(m->glx, T->glx, nc); cp(m->grx, T->grx, nc);
    cp(m->efx, T->efx, nc); cp(m->cwx, T->cwx, nc); cp(m->ftabx, T->ftabx, nc * 2 * T->rows);
    cp(m->fly, T->fly, nc); cp(m->fry, T->fry, nc); cp(m->gly, T->gly, nc); cp(m->gry, T->gry, nc);
    cp(m->efy, T->efy, nc); cp(m->cwy, T->cwy, nc); cp(m->ftaby, T->ftaby, nc * 2 * T->rows);
    m->build_exchange(T->nbr);
    // static_cell_state (state.py:216-223): own entries from the tiles, halos by the exchange of every tile
    m->zcx.assign(n, R(0));
    m->zcy.assign(n, R(0));
    for (int t = 1; t < T->ntiles; ++t)
        for (int j = 0; j < TILE; ++j)
            for (int i = 0; i < TILE; ++i) {
                const size_t k = (size_t)t * STRIDE2 + (j + HALO) * STRIDE + i + HALO;
                m->zcx[k] = R(T->zcx[(size_t)t * TILE2 + j * TILE + i]);
                m->zcy[k] = R(T->zcy[(size_t)t * TILE2 + j * TILE + i]);
            }
    for (int t = 1; t < T->ntiles; ++t)
        for (int q = 0; q < m->nhalo; ++q) {
            const size_t dst = (size_t)t * STRIDE2 + m->slot[q];
            const size_t s = m->src[(size_t)(t - 1) * m->nhalo + q];
            m->zcx[dst] = m->zcx[s];
            m->zcy[dst] = m->zcy[s];
        }
    m->wave.assign(T->ntiles, R(0));
    m->wet.assign(T->ntiles, 0);
    for (int t = 1; t < T->ntiles; ++t) m->active.push_back(t);
    return m;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// C entry points for oracle/oracle.py (ctypes). `prec` is 8 (double, the FP64 oracle) or 4 (float, the baseline).
// Fields are exchanged as double on the Python side whatever the model's type.

struct OracleHandle {
    int prec;
    void* m;
};

template <class R>
static Model<R>* M(OracleHandle* h) { return static_cast<Model<R>*>(h->m); }

#define DISPATCH(h, call) ((h)->prec == 8 ? (M<double>(h)->call) : (M<float>(h)->call))

extern "C" {

OracleHandle* oracle_create(const Tables* T, const Params* P, int prec, int threads) {
    auto* h = new OracleHandle{prec, nullptr};
    if (prec == 8)
        h->m = create<double>(T, P, threads);
    else
        h->m = create<float>(T, P, threads);
    return h;
}

void oracle_destroy(OracleHandle* h) {
    if (!h) return;
    if (h->prec == 8)
        delete M<double>(h);
    else
        delete M<float>(h);
    delete h;
}

int oracle_field_count() { return F_COUNT; }

void oracle_set_field(OracleHandle* h, int f, const double* v, int64_t n) {
    if (h->prec == 8) {
        auto& a = M<double>(h)->a[f];
        for (int64_t i = 0; i < n; ++i) a[i] = v[i];
    } else {
        auto& a = M<float>(h)->a[f];
        for (int64_t i = 0; i < n; ++i) a[i] = (float)v[i];
    }
}

void oracle_get_field(OracleHandle* h, int f, double* v, int64_t n) {
    if (h->prec == 8) {
        auto& a = M<double>(h)->a[f];
        for (int64_t i = 0; i < n; ++i) v[i] = a[i];
    } else {
        auto& a = M<float>(h)->a[f];
        for (int64_t i = 0; i < n; ++i) v[i] = a[i];
    }
}

void oracle_set_active(OracleHandle* h, const int32_t* tiles, int n) {
    if (h->prec == 8)
        M<double>(h)->active.assign(tiles, tiles + n);
    else
        M<float>(h)->active.assign(tiles, tiles + n);
}

void oracle_set_boundary(OracleHandle* h, const int32_t* cell, const int32_t* ref, const int32_t* kind,
                         const float* d0, const float* h0, int n) {
    auto set = [&](auto* m) {
        m->bcell.assign(cell, cell + n);
        m->bref.assign(ref, ref + n);
        m->bkind.assign(kind, kind + n);
        m->bd0.assign(d0, d0 + n);
        m->bh0.assign(h0, h0 + n);
    };
    if (h->prec == 8)
        set(M<double>(h));
    else
        set(M<float>(h));
}

void oracle_step(OracleHandle* h, double dt, int want_wave, int want_wet) {
    DISPATCH(h, step(dt, want_wave != 0, want_wet != 0));
}

double oracle_max_wave(OracleHandle* h) {
    double w = 0.0;
    if (h->prec == 8) {
        for (double v : M<double>(h)->wave) w = std::max(w, v);
    } else {
        float wf = 0.0f;
        for (float v : M<float>(h)->wave) wf = std::max(wf, v);
        w = wf;
    }
    return w;
}

void oracle_wave(OracleHandle* h, double* out) {
    if (h->prec == 8) {
        auto& w = M<double>(h)->wave;
        std::copy(w.begin(), w.end(), out);
    } else {
        auto& w = M<float>(h)->wave;
        for (size_t i = 0; i < w.size(); ++i) out[i] = w[i];
    }
}

void oracle_wet_flags(OracleHandle* h, int32_t* out) {
    if (h->prec == 8) {
        auto& w = M<double>(h)->wet;
        std::copy(w.begin(), w.end(), out);
    } else {
        auto& w = M<float>(h)->wet;
        std::copy(w.begin(), w.end(), out);
    }
}

void oracle_set_wet(OracleHandle* h, const int32_t* flags) {
    if (h->prec == 8) {
        auto& w = M<double>(h)->wet;
        std::copy(flags, flags + w.size(), w.begin());
    } else {
        auto& w = M<float>(h)->wet;
        std::copy(flags, flags + w.size(), w.begin());
    }
}

void oracle_refresh_plan(OracleHandle* h, const int32_t* tiles, int n) { DISPATCH(h, plan(tiles, n)); }

void oracle_refresh_halos(OracleHandle* h) { DISPATCH(h, refresh_halos()); }

// ---- local time stepping (cpu/levels.py CpuLevels): one call per operation of the level stepper
void oracle_lts_init(OracleHandle* h) { DISPATCH(h, lts_init()); }

void oracle_lts_set_level(OracleHandle* h, const int32_t* level) {
    auto set = [&](auto* m) { std::copy(level, level + m->ntiles, m->level.begin()); };
    if (h->prec == 8)
        set(M<double>(h));
    else
        set(M<float>(h));
}

// op: 0 halo_state (arg n, flag want_plan), 1 faces, 2 faces_done, 3 limit, 4 limit_done, 5 save, 6 update (flag
// want_plan), 7 commit, 8 clear_wave, 9 clear_wet, 10 boundary, 11 refresh_plan
void oracle_lts_tiles(OracleHandle* h, int op, const int32_t* tiles, int n, double dt, int flag, int arg) {
    auto run = [&](auto* m) {
        using R = typename std::remove_pointer<decltype(m)>::type::Real;
        const R d = R(dt);
        switch (op) {
        case 0:
            m->halo_levels(tiles, n, arg);
            if (flag) m->exchange_list(tiles, n, m->a[F_OM].data());
            break;
        case 1: m->faces_list(tiles, n, d, flag != 0); break;
        case 2:
            for (Field f : {F_U1, F_V1, F_QX1, F_QY1}) m->exchange_list(tiles, n, m->a[f].data());
            break;
        case 3: m->limit_list(tiles, n, d); break;
        case 4: m->exchange_list(tiles, n, m->a[F_R].data()); break;
        case 5: m->save_old(tiles, n); break;
        case 6: m->update_list(tiles, n, d, flag != 0, true); break;
        case 7: m->commit(tiles, n); break;
        case 8: std::fill(m->wave.begin(), m->wave.end(), R(0)); break;
        case 9: std::fill(m->wet.begin(), m->wet.end(), 0); break;
        case 10:
            if (!m->bcell.empty()) m->boundary();
            break;
        case 11: m->plan(tiles, n); break;
        }
    };
    if (h->prec == 8)
        run(M<double>(h));
    else
        run(M<float>(h));
}

// op: 0 strips (flag want_wave), 1 commit_strips, 2 estimate, 3 flux; `edges` rows of 4
void oracle_lts_edges(OracleHandle* h, int op, const int32_t* edges, int n, double dt, int flag) {
    auto run = [&](auto* m) {
        using R = typename std::remove_pointer<decltype(m)>::type::Real;
        const R d = R(dt);
        switch (op) {
        case 0: m->strips(edges, n, d, flag != 0); break;
        case 1: m->commit_strips(edges, n); break;
        case 2: m->edge_estimate(edges, n); break;
        case 3: m->edge_flux(edges, n, d); break;
        }
    };
    if (h->prec == 8)
        run(M<double>(h));
    else
        run(M<float>(h));
}

void oracle_set_threads(OracleHandle* h, int threads) {
    if (h->prec == 8)
        M<double>(h)->threads = std::max(threads, 1);
    else
        M<float>(h)->threads = std::max(threads, 1);
}

}  // extern "C"
