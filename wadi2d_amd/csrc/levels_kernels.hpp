// Local time stepping on the device (behaviour of wadi2d/solver/gpu/kernels/levels.cu:1-172 and the tile-list forms
// of faces.cu, cells.cu). Included inside step.hip's kernel namespace: it uses that file's face(), staging,
// Bufs/Cur and table lookups.
//
// A tile at level L steps 2^L smallest steps at a time. The arrays stay the uniform step's, at one fixed parity for
// the whole level run: h, u, v, mx, my are the current buffers (h updated in place, as the limit is a launch of its
// own), u1, v1, mx1, my1 the other buffers, committed tile by tile. There are no halos: a tile reads its neighbours'
// own cells, so "exchanging" a halo is nothing, and the one thing a halo did beyond a copy, reading a neighbour that
// is in the middle of its step at a time between its old and new values, is done where the neighbour's depth and
// level are staged.
//
// An edge row is (fine tile, side of the fine tile: 0 west 1 north 2 east 3 south, coarse tile, edge number).

template <class R>
struct LevelArgs {
    StepArgs<R> A;          // buffers, statics, parameters; A.par is the level run's parity
    const int* list;        // tiles, or edge rows of 4
    int n;
    R dt;
    int flag;               // want_wave (faces, strips), want_plan (update)
    const int* level;       // per tile
    const R* h0;            // own layout: depth and level of a tile before its current step
    const R* d0;
    int sub;                // the smallest step within the cycle (halo time)
    R* r;                   // own layout: the positivity ratio
    R* est;                 // per edge and face: the estimate taken when the coarse step began
};

// The depth and level of ring entries whose tile is in the middle of its step at smallest step `sub`: between its
// values before and after the step (halo_levels, levels.cu:67-91). Each thread rewrites what it staged.
template <class R, int NTH>
__device__ __forceinline__ void ring_in_time(const Stager<NTH>& sg, R* s_h, R* s_d, const R* h, const R* d,
                                             const LevelArgs<R>& L) {
#pragma unroll
    for (int m = 0; m < Stager<NTH>::RING; ++m) {
        if (sg.ring_l[m] < 0) continue;
        const int g = sg.ring_g[m];
        const int span = 1 << L.level[g / TILE2];
        const int part = L.sub & (span - 1);
        if (part == 0) continue;
        const R w = R(part) / R(span);
        s_h[sg.ring_l[m]] = L.h0[g] + w * (h[g] - L.h0[g]);
        s_d[sg.ring_l[m]] = L.d0[g] + w * (d[g] - L.d0[g]);
    }
}

template <class R>
__device__ __forceinline__ R* other_of(R* const (&b)[2], int par) { return (par & 1) ? b[0] : b[1]; }

// The six staged arrays of tile t (depth and level read in time), for faces_level_kernel and strips_kernel.
template <class R, int NTH>
__device__ __forceinline__ void stage_level(const LevelArgs<R>& L, const Cur<R>& s, int t, int* nb, R* s_u, R* s_v,
                                            R* s_h, R* s_d, R* s_qx, R* s_qy) {
    load_nb3(t, L.A.st.nbr, nb);
    __syncthreads();
    const Stager<NTH> sg = make_stager<NTH>(t, nb);
    R* const dst[6] = {s_u, s_v, s_h, s_d, s_qx, s_qy};
    const R* const src[6] = {s.u, s.v, s.h, s.d, s.qx, s.qy};
    stage<R, 6, NTH>(dst, src, sg);
    ring_in_time<R, NTH>(sg, s_h, s_d, s.h, s.d, L);
    __syncthreads();
}

// One own face of cell (j, i) of tile t on axis x (west face) or y (north face), from the staged block; writes the
// face's velocity, momentum and discharge into the step's next buffers and returns its wave speed.
template <class R, bool FAST>
__device__ __forceinline__ R level_face(const LevelArgs<R>& L, const Cur<R>& s, R* mx1, R* my1, int t, int j, int i,
                                        bool xaxis, const int* nb, const FaceIn<R>& fin, bool want_wave) {
    const StepArgs<R>& A = L.A;
    const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
    const int Lx = lds_of(j, i);
    const float2 e = A.st.es[c];
    R un, mn, qn, w;
    if (xaxis) {
        const R ef = R(fabsf(e.x));
        const R cw = signbit(e.x) ? R(A.st.cwv[c].x) : R(0);
        const unsigned k = west_of(t, j, i, nb);
        w = face<R, FAST>(fin, Lx, s.mx[c], A.st.sx[c], ef, cw, R(A.st.zcx[k]), R(A.st.zcx[c]), A.st.ftabx + c,
                          A.st.ncell, A.st.last, A.st.fcx ? A.st.fcx + c : nullptr, A.st.ftagx + c,
                          want_wave ? s.om[k] : R(0), want_wave ? s.om[c] : R(0), L.dt, want_wave, A.P, un, mn, qn);
        s.u1[c] = un;
        mx1[c] = mn;
        s.qx1[c] = qn;
    } else {
        const R ef = R(fabsf(e.y));
        const R cw = signbit(e.y) ? R(A.st.cwv[c].y) : R(0);
        const unsigned k = north_of(t, j, i, nb);
        w = face<R, FAST>(fin, Lx, s.my[c], A.st.sy[c], ef, cw, R(A.st.zcy[k]), R(A.st.zcy[c]), A.st.ftaby + c,
                          A.st.ncell, A.st.last, A.st.fcy ? A.st.fcy + c : nullptr, A.st.ftagy + c,
                          want_wave ? s.om[k] : R(0), want_wave ? s.om[c] : R(0), L.dt, want_wave, A.P, un, mn, qn);
        s.v1[c] = un;
        my1[c] = mn;
        s.qy1[c] = qn;
    }
    return w;
}

// Both own faces of every cell of the listed tiles at one dt (faces.cu:244-271 over a level's tiles); the tile's
// wave speed is set, not raised, as the uniform step sets it.
template <class R, bool FAST>
__global__ __launch_bounds__(NT) void faces_level_kernel(LevelArgs<R> L) {
    __shared__ __align__(16) R s_u[LSIZE];
    __shared__ __align__(16) R s_v[LSIZE];
    __shared__ __align__(16) R s_h[LSIZE];
    __shared__ __align__(16) R s_d[LSIZE];
    __shared__ __align__(16) R s_qx[LSIZE];
    __shared__ __align__(16) R s_qy[LSIZE];
    __shared__ int nb[9];
    __shared__ R red[NT / 64];
    const int t = uni(L.list[blockIdx.x]);
    const Cur<R> s = view(L.A.b, L.A.par);
    R* mx1 = other_of(L.A.b.mx, L.A.par);
    R* my1 = other_of(L.A.b.my, L.A.par);
    stage_level<R, NT>(L, s, t, nb, s_u, s_v, s_h, s_d, s_qx, s_qy);
    const FaceIn<R> fx{s_u, s_v, s_qx, s_qy, s_h, s_d, 1, LS};
    const FaceIn<R> fy{s_v, s_u, s_qy, s_qx, s_h, s_d, LS, 1};
    const bool want = L.flag != 0;
    const int i = threadIdx.x & (TILE - 1);
    R w = R(0);
    for (int r = 0; r < TILE2 / NT; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NT / TILE);
        const R wx = level_face<R, FAST>(L, s, mx1, my1, t, j, i, true, nb, fx, want);
        const R wy = level_face<R, FAST>(L, s, mx1, my1, t, j, i, false, nb, fy, want);
        const R m = wx > wy ? wx : wy;
        w = m > w ? m : w;
    }
    if (want) {
        const R m = block_max<NT>(w, red);
        if (threadIdx.x == 0) L.A.wave[t] = m;
    }
}

// A wave speed raised in place by an atomic max of its bits (wave speeds are never negative).
__device__ __forceinline__ void raise_wave(float* p, float v) { atomicMax((unsigned*)p, __float_as_uint(v)); }
__device__ __forceinline__ void raise_wave(double* p, double v) {
    atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

// The faces of an edge that the coarse tile owns (the fine tile's east or south side), stepped at the fine rate
// (strips, levels.cu:34-64): one block per row, stages the coarse tile, one thread per face; the wave speed raises
// both tiles'.
template <class R, bool FAST>
__global__ __launch_bounds__(NT) void strips_kernel(LevelArgs<R> L) {
    __shared__ __align__(16) R s_u[LSIZE];
    __shared__ __align__(16) R s_v[LSIZE];
    __shared__ __align__(16) R s_h[LSIZE];
    __shared__ __align__(16) R s_d[LSIZE];
    __shared__ __align__(16) R s_qx[LSIZE];
    __shared__ __align__(16) R s_qy[LSIZE];
    __shared__ int nb[9];
    const int* e = L.list + 4 * blockIdx.x;
    const int fine = uni(e[0]), side = uni(e[1]), t = uni(e[2]);
    const Cur<R> s = view(L.A.b, L.A.par);
    R* mx1 = other_of(L.A.b.mx, L.A.par);
    R* my1 = other_of(L.A.b.my, L.A.par);
    stage_level<R, NT>(L, s, t, nb, s_u, s_v, s_h, s_d, s_qx, s_qy);
    if (threadIdx.x >= TILE) return;
    const int q = threadIdx.x;
    const bool want = L.flag != 0;
    R w;
    if (side == 2) {
        const FaceIn<R> fx{s_u, s_v, s_qx, s_qy, s_h, s_d, 1, LS};
        w = level_face<R, FAST>(L, s, mx1, my1, t, q, 0, true, nb, fx, want);
    } else {
        const FaceIn<R> fy{s_v, s_u, s_qy, s_qx, s_h, s_d, LS, 1};
        w = level_face<R, FAST>(L, s, mx1, my1, t, 0, q, false, nb, fy, want);
    }
    if (want) {
        for (int off = 16; off > 0; off >>= 1) {
            const R o = __shfl_xor(w, off, 32);
            w = o > w ? o : w;
        }
        if (q == 0) {
            raise_wave(L.A.wave + t, w);
            raise_wave(L.A.wave + fine, w);
        }
    }
}

// State index of the cell (j, i) of tile t, with j or i one step past the tile read from the neighbour there.
__device__ __forceinline__ unsigned east_of(int t, int j, int i, const int* __restrict__ nbr) {
    return i < TILE - 1 ? t * TILE2 + j * TILE + i + 1 : nbr[(size_t)t * 8 + 4] * TILE2 + j * TILE;
}
__device__ __forceinline__ unsigned south_of(int t, int j, int i, const int* __restrict__ nbr) {
    return j < TILE - 1 ? t * TILE2 + (j + 1) * TILE + i : nbr[(size_t)t * 8 + 6] * TILE2 + i;
}
__device__ __forceinline__ unsigned west_cell(int t, int j, int i, const int* __restrict__ nbr) {
    return i > 0 ? t * TILE2 + j * TILE + i - 1 : nbr[(size_t)t * 8 + 3] * TILE2 + j * TILE + TILE - 1;
}
__device__ __forceinline__ unsigned north_cell(int t, int j, int i, const int* __restrict__ nbr) {
    return j > 0 ? t * TILE2 + (j - 1) * TILE + i : nbr[(size_t)t * 8 + 1] * TILE2 + (TILE - 1) * TILE + i;
}

// The positivity ratio of the listed tiles' cells at one dt (limit, cells.cu:4-18).
template <class R, bool FAST>
__global__ __launch_bounds__(NT) void limit_level_kernel(LevelArgs<R> L) {
    const int t = uni(L.list[blockIdx.x]);
    const Cur<R> s = view(L.A.b, L.A.par);
    const int* nbr = L.A.st.nbr;
    const R dx = R(L.A.P.dx);
    const R area = dx * dx;
    const int i = threadIdx.x & (TILE - 1);
    for (int r = 0; r < TILE2 / NT; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NT / TILE);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        const R out = pmax(s.qx1[east_of(t, j, i, nbr)], R(0)) + pmax(-s.qx1[c], R(0)) +
                      pmax(s.qy1[south_of(t, j, i, nbr)], R(0)) + pmax(-s.qy1[c], R(0));
        const R have = pmax(s.h[c], R(0)) * area;
        L.r[c] = out * L.dt > have ? dv<FAST>(have, out * L.dt) : R(1);
    }
}

// Continuity, rain, the residual (an overdrawn cell keeps what it owes), level and wet share of the listed tiles at
// one dt, the depth updated in place (update, cells.cu:28-63 with debt); every listed tile's wet flag is written.
template <class R, bool FAST>
__global__ __launch_bounds__(NT) void update_level_kernel(LevelArgs<R> L) {
    const int t = uni(L.list[blockIdx.x]);
    const StepArgs<R>& A = L.A;
    const Cur<R> s = view(A.b, A.par);
    const int* nbr = A.st.nbr;
    const R dx = R(A.P.dx);
    const R area = dx * dx;
    const R dt = L.dt;
    const bool want_plan = L.flag != 0;
    const int i = threadIdx.x & (TILE - 1);
    bool wet = false;
    for (int r = 0; r < TILE2 / NT; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NT / TILE);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        const unsigned ce = east_of(t, j, i, nbr), cs = south_of(t, j, i, nbr);
        const R rc = L.r[c];
        const R qw = scaled(s.qx1[c], L.r[west_cell(t, j, i, nbr)], rc);
        const R qe = scaled(s.qx1[ce], rc, L.r[ce]);
        const R qn = scaled(s.qy1[c], L.r[north_cell(t, j, i, nbr)], rc);
        const R qs = scaled(s.qy1[cs], rc, L.r[cs]);
        s.qx[c] = qw;
        s.qy[c] = qn;
        const R h0 = s.h[c];
        const R inc = s.res[c] + dt * s.src[c] - dv<FAST>(dt * (qe - qw + qs - qn), area);
        const R sum = h0 + inc;
        const R hn = pmax(sum, R(0));
        s.res[c] = sum > R(0) ? inc - (sum - h0) : sum;
        s.h[c] = hn;
        const R Hc = R(A.st.hc[c]);
        s.d[c] = cell_level<R, FAST>(hn, Hc, A.st.ctab + c, A.st.ncell, A.st.last);
        if (want_plan) s.om[c] = cell_plan<R, FAST>(hn, Hc, A.st.ctab + c, A.st.ncell, A.st.last);
        wet |= hn > R(0);
    }
    const int any = __syncthreads_or(wet ? 1 : 0);
    if (threadIdx.x == 0) A.wet[t] = any;
}

// The old depth and level of the listed tiles, for finer neighbours reading them in time (save_old,
// levels.cu:93-105); the listed tiles' faces into the current buffers (commit, levels.cu:107-122).
template <class R>
__global__ __launch_bounds__(NT) void save_level_kernel(LevelArgs<R> L, R* h0, R* d0) {
    const int t = uni(L.list[blockIdx.x]);
    const Cur<R> s = view(L.A.b, L.A.par);
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const unsigned c = (unsigned)(t * TILE2 + p);
        h0[c] = s.h[c];
        d0[c] = s.d[c];
    }
}

template <class R>
__global__ __launch_bounds__(NT) void commit_level_kernel(LevelArgs<R> L) {
    const int t = uni(L.list[blockIdx.x]);
    const Cur<R> s = view(L.A.b, L.A.par);
    const R* mx1 = other_of(L.A.b.mx, L.A.par);
    const R* my1 = other_of(L.A.b.my, L.A.par);
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const unsigned c = (unsigned)(t * TILE2 + p);
        s.u[c] = s.u1[c];
        s.v[c] = s.v1[c];
        s.mx[c] = mx1[c];
        s.my[c] = my1[c];
    }
}

// The edge faces the coarse tile owns, committed (commit_strips, levels.cu:124-140): 32 threads per row.
template <class R>
__global__ __launch_bounds__(TILE) void commit_strips_kernel(LevelArgs<R> L) {
    const int* e = L.list + 4 * blockIdx.x;
    const int side = e[1], t = e[2], q = threadIdx.x;
    const Cur<R> s = view(L.A.b, L.A.par);
    if (side == 2) {
        const unsigned c = (unsigned)(t * TILE2 + q * TILE);
        s.u[c] = s.u1[c];
        s.mx[c] = other_of(L.A.b.mx, L.A.par)[c];
    } else {
        const unsigned c = (unsigned)(t * TILE2 + q);
        s.v[c] = s.v1[c];
        s.my[c] = other_of(L.A.b.my, L.A.par)[c];
    }
}

// An edge's face q: the face (state index and axis), the cells on its low and high side, and the coarse cell whose
// residual takes the corrections (edge_face and edge_cell, levels.cu:12-31, in the own layout).
struct EdgeFace {
    unsigned face, low, high, coarse;
    bool y;
};
__device__ __forceinline__ EdgeFace edge_face_of(int fine, int side, int crs, int q) {
    EdgeFace f;
    const unsigned F = (unsigned)fine * TILE2, C = (unsigned)crs * TILE2;
    if (side == 0) {
        f = {F + q * TILE, C + q * TILE + TILE - 1, F + q * TILE, C + q * TILE + TILE - 1, false};
    } else if (side == 1) {
        f = {F + q, C + (TILE - 1) * TILE + q, F + q, C + (TILE - 1) * TILE + q, true};
    } else if (side == 2) {
        f = {C + q * TILE, F + q * TILE + TILE - 1, C + q * TILE, C + q * TILE, false};
    } else {
        f = {C + q, F + (TILE - 1) * TILE + q, C + q, C + q, true};
    }
    return f;
}

// The edge's discharge as both tiles apply it when the coarse step starts (edge_estimate, levels.cu:143-154).
template <class R>
__global__ __launch_bounds__(TILE) void estimate_kernel(LevelArgs<R> L) {
    const int* e = L.list + 4 * blockIdx.x;
    const int q = threadIdx.x;
    const EdgeFace f = edge_face_of(e[0], e[1], e[2], q);
    const Cur<R> s = view(L.A.b, L.A.par);
    const R flow = f.y ? s.qy1[f.face] : s.qx1[f.face];
    L.est[(size_t)e[3] * TILE + q] = scaled(flow, L.r[f.low], L.r[f.high]);
}

// A later fine step: the coarse cell takes what the edge carried beyond the estimate into its residual; a coarse
// tile that owns the face keeps its limited discharge (edge_flux, levels.cu:156-172). One axis per launch, so no
// cell is written twice.
template <class R, bool FAST>
__global__ __launch_bounds__(TILE) void flux_kernel(LevelArgs<R> L) {
    const int* e = L.list + 4 * blockIdx.x;
    const int side = e[1], q = threadIdx.x;
    const EdgeFace f = edge_face_of(e[0], side, e[2], q);
    const Cur<R> s = view(L.A.b, L.A.par);
    const R flow = f.y ? s.qy1[f.face] : s.qx1[f.face];
    const R qs = scaled(flow, L.r[f.low], L.r[f.high]);
    const R sign = side < 2 ? R(-1) : R(1);
    const R dx = R(L.A.P.dx);
    s.res[f.coarse] += dv<FAST>(sign * L.dt * (qs - L.est[(size_t)e[3] * TILE + q]), dx * dx);
    if (side == 2) s.qx[f.face] = qs;
    if (side == 3) s.qy[f.face] = qs;
}
