// The uniform-grid shallow-water step on CDNA4 (gfx950), and its C-ABI (include/wadi2d_hip.h).
//
// Scheme: Wadi2D's sub-grid finite-volume step (wadi2d/solver/gpu/kernels/faces.cu, cells.cu, common.cuh; the
// statement order of its Numba twin wadi2d/solver/cpu/kernels.py). The design here is this library's own:
//
//  * Device layout. Tiles of 32 x 32 own cells, tile-major, and no halo entries: cell (t, j, i) is at
//    t * 1024 + j * 32 + i in every state array. Tile 0 is the void tile, all zero and never written. A kernel
//    reads the 2-cell ring around its tile straight from the 8 neighbour tiles into LDS, so no exchange kernels
//    and no halo traffic in HBM beyond the ring reads themselves.
//  * Two kernels per step. `faces_kernel` computes both own faces of every own cell (velocity, momentum,
//    discharge) and the tile's wave speed. `cells_kernel` recomputes the positivity ratio r on the tile and a
//    one-cell ring in LDS (the reference's `limit` kernel and its exchange) and applies continuity, rain and the
//    compensated residual, the level table inversion, the wet share and the tile's wet flag.
//  * Tables. Face tables are plane-major records of consecutive row pairs, float4 (A_r, T_r, A_r+1, T_r+1) at
//    r * ncell + cell, so a lookup is one 16-byte load and neighbouring lanes at the same row read one run of
//    memory. Cell tables likewise, float2 (d_r, d_r+1). Face statics are packed float4 (fl, fr, gl, gr) per
//    direction and (efx, cwx, efy, cwy).
//  * Mean depth is double-buffered: the cells kernel reads the ring of a neighbour tile that another workgroup
//    updates in the same launch.
//
// Arithmetic variants: R = double (the FP64 path), R = float exact (IEEE division/sqrt, cube root rounded once
// from double, no contraction: the reference's float32 statement order), R = float fast (hardware rcp/sqrt and
// a log2/exp2 cube root).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mlp.hpp"
#include "wadi2d_hip.h"

#pragma clang fp contract(off)

namespace {

constexpr int TILE = 32;
constexpr int TILE2 = TILE * TILE;
constexpr int HALO = 2;
constexpr int STRIDE = TILE + 2 * HALO;  // 36: the reference's state row, and the LDS row here
constexpr int STRIDE2 = STRIDE * STRIDE;
constexpr int NT = 256;                   // threads per workgroup: one tile
constexpr int ROWS_PER_THREAD = TILE2 / NT;  // the cell kernel's

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? 2 : 3, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------------------------
// arithmetic

template <bool FAST>
__device__ __forceinline__ float dv(float a, float b) {
    if constexpr (FAST) return a * __builtin_amdgcn_rcpf(b);
    return a / b;
}
template <bool FAST>
__device__ __forceinline__ double dv(double a, double b) { return a / b; }

template <bool FAST>
__device__ __forceinline__ float sq(float x) {
    if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
    return sqrtf(x);
}
template <bool FAST>
__device__ __forceinline__ double sq(double x) { return sqrt(x); }

template <bool FAST>
__device__ __forceinline__ float cb(float x) {
    if constexpr (FAST) return x > 0.0f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f)) : 0.0f;
    return (float)cbrt((double)x);
}
template <bool FAST>
__device__ __forceinline__ double cb(double x) { return cbrt(x); }

template <class R>
__device__ __forceinline__ R maxr(R a, R b) { return a > b ? a : (b > a ? b : a); }  // np.maximum-free: as max()
template <class R>
__device__ __forceinline__ R minr(R a, R b) { return a < b ? a : (b < a ? b : a); }

// Python's max(a, b) returns a unless b > a; min(a, b) returns a unless b < a. The ternaries below keep that
// order for the NaN cases the reference's float32 code can meet.
template <class R>
__device__ __forceinline__ R pmax(R a, R b) { return b > a ? b : a; }
template <class R>
__device__ __forceinline__ R pmin(R a, R b) { return b < a ? b : a; }

// max/min of the step: Python's order in the exact variants, the hardware's v_max/v_min (which differ only for a
// NaN and the sign of a zero) in the fast one.
template <bool FAST, class R>
__device__ __forceinline__ R fmx(R a, R b) {
    if constexpr (FAST) return fmax(a, b);
    return pmax(a, b);
}
template <bool FAST, class R>
__device__ __forceinline__ R fmn(R a, R b) {
    if constexpr (FAST) return fmin(a, b);
    return pmin(a, b);
}

// int(s) for the row of a table, capped; NaN and negative give row 0.
template <class R>
__device__ __forceinline__ int row_of(R s, int cap) {
    if (!(s > R(0))) return 0;
    return s < R(cap) ? (int)s : cap;
}

template <class R>
__device__ __forceinline__ R minmod(R a, R b) {
    if (a * b <= R(0)) return R(0);
    return fabs(a) < fabs(b) ? a : b;
}

template <class R>
__device__ __forceinline__ R ustar(R uc, R uup, R udown, bool second) {
    if (!second) return uc;
    return uc + R(0.5) * minmod(uc - uup, udown - uc);
}

template <class R>
__device__ __forceinline__ R gated(R n, R own, bool mirror) { return (mirror && n == R(0)) ? own : n; }

// Streams a step touches once (per-cell statics, row caches, the kernels' outputs) go through nontemporal loads and
// stores, so that they do not push out of L2 the arrays whose 2-cell rings neighbouring workgroups read again.
// Measured on the 4096 x 4096 step: face kernel 0.534 -> 0.496 ms per launch.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float4 ld_nt(const float4* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    typedef float v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_float2(v.x, v.y);
}
template <class T>
__device__ __forceinline__ void st_nt(T v, T* p) { __builtin_nontemporal_store(v, p); }

struct Params {
    double g, dx, e_wet, h_cfl, u_max, lead, a_min;
    int rows, second, sheet, mirror, carry;
};

// The row pair `r` of a table record, through the face's row cache when it has one: the cache holds the pair of
// the row last read and its row in `tag`; both were read (`pre`, `have`) before the row was known, so a lookup whose
// row has not changed since the last step waits for no further load; a row change reads the table and refills it.
template <class V>
__device__ __forceinline__ V fetch_row(const V* __restrict__ tab, size_t plane, int r, V* cache, unsigned char* tag,
                                       int have, V pre) {
    if (cache) {
        if (have == r) return pre;
        const V q = tab[(size_t)r * plane];
        *cache = q;
        *tag = (unsigned char)r;
        return q;
    }
    return tab[(size_t)r * plane];
}

// Area, conveyance^(3/5), wet width at depth e over the invert (common.cuh:64-90): rows quadratic in e, read
// linearly between them, prismatic above the table top E. `tab` points at this face in plane 0; planes are
// `plane` records apart. `cache`/`tag`/`have`/`pre`: the face's row cache (null for none), its tag and pair as read.
template <class R, bool FAST>
__device__ __forceinline__ void face_lookup(R e, R E, const float4* __restrict__ tab, size_t plane, int last, R& A,
                                            R& T, R& W, float4* cache = nullptr, unsigned char* tag = nullptr,
                                            int have = -1, float4 pre = float4{0.f, 0.f, 0.f, 0.f}) {
    const R m = R(last);
    if (e >= E) {
        const float4 q = fetch_row(tab, plane, last - 1, cache, tag, have, pre);
        const R a1 = R(q.z), a0 = R(q.x);
        const R c = dv<FAST>(m - R(1), m);
        W = dv<FAST>(a1 - a0, E * (R(1) - c * c));
        A = a1 + W * (e - E);
        T = R(q.w) * dv<FAST>(A, pmax(a1, R(1e-12)));
        return;
    }
    if constexpr (FAST) {
        // the same rows, fewer divisions: depth over the top read once as a ratio, one reciprocal per row gap
        const R x = e * __builtin_amdgcn_rcpf(E);
        const R s = sq<FAST>(x) * m;
        const int r = row_of(s, last - 1);
        const float4 q = fetch_row(tab, plane, r, cache, tag, have, pre);
        const R im = __builtin_amdgcn_rcpf(m);
        const R c0 = R(r) * im;
        const R c1 = R(r + 1) * im;
        const R c0s = c0 * c0;
        const R gap = c1 * c1 - c0s;
        const R ig = __builtin_amdgcn_rcpf(gap);
        const R f = (x - c0s) * ig;
        const R a0 = R(q.x), a1 = R(q.z), t0 = R(q.y);
        A = a0 + f * (a1 - a0);
        T = t0 + f * (R(q.w) - t0);
        W = (a1 - a0) * ig * __builtin_amdgcn_rcpf(E);
        return;
    }
    const R s = sq<FAST>(dv<FAST>(e, E)) * m;
    const int r = row_of(s, last - 1);
    const float4 q = fetch_row(tab, plane, r, cache, tag, have, pre);
    const R c0 = dv<FAST>(R(r), m);
    const R c1 = dv<FAST>(R(r + 1), m);
    const R e0 = c0 * c0 * E;
    const R de = (c1 * c1 - c0 * c0) * E;
    const R f = dv<FAST>(e - e0, de);
    const R a0 = R(q.x), a1 = R(q.z), t0 = R(q.y);
    A = a0 + f * (a1 - a0);
    T = t0 + f * (R(q.w) - t0);
    W = dv<FAST>(a1 - a0, de);
}

// Level above zmin from the mean depth, and the wet share of the plan (common.cuh:39-56).
template <class R, bool FAST>
__device__ __forceinline__ R cell_level(R hb, R Hc, const float2* __restrict__ tab, size_t plane, int last) {
    if (Hc <= R(0)) return hb;
    if (hb >= Hc) return R(tab[(size_t)(last - 1) * plane].y) + (hb - Hc);
    const R s = sq<FAST>(dv<FAST>(hb, Hc)) * R(last);
    const int r = row_of(s, last - 1);
    const float2 q = tab[(size_t)r * plane];
    return R(q.x) + (s - R(r)) * (R(q.y) - R(q.x));
}

template <class R, bool FAST>
__device__ __forceinline__ R cell_plan(R hb, R Hc, const float2* __restrict__ tab, size_t plane, int last) {
    if (Hc <= R(0) || hb >= Hc) return R(1);
    const R s = sq<FAST>(dv<FAST>(hb, Hc)) * R(last);
    const int r = row_of(s, last - 1);
    const float2 q = tab[(size_t)r * plane];
    const R m = R(last);
    return dv<FAST>(dv<FAST>(Hc * R(2 * r + 1), m * m), pmax(R(q.y) - R(q.x), R(1e-12)));
}

// ------------------------------------------------------------------------------------------------------------
// device arrays of one model

// Every state buffer of one model. Mean depth and the face arrays have two buffers, the current one and the next:
// the cell kernel reads the ring of a neighbour tile that another workgroup updates in the same launch, and the face
// kernel writes the face velocity (and momentum) the next step reads while this step's are still read. The face
// kernel reads level, discharge and wet share before the cell kernel rewrites them, so those are updated in place.
// `qx1, qy1` hold the face discharges before the positivity limit, between the two kernels. The residual and the
// discharges before the limit have two buffers as well: the pipelined step recomputes a neighbour tile's ring from its
// old residual and discharges while that tile writes its new ones.
template <class R>
struct Bufs {
    R *h[2], *u[2], *v[2], *mx[2], *my[2], *res[2], *qx1[2], *qy1[2];
    R *d, *qx, *qy, *om, *src;
};

// A value every lane of the wave holds alike, kept in a scalar register (the compiler cannot see that a value loaded
// from a tile's member is uniform, and would otherwise carry every buffer pointer chosen from it in vector registers).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// One step's view of the buffers at parity `par` (the steps taken, mod 2): the current and the next buffer of each
// buffered array (the next mean depth is also the one the last step read, from which the face kernel recomputes the
// momentum before the cell kernel overwrites it).
template <class R>
struct Cur {
    R *h, *h_next, *u, *v, *u1, *v1, *mx, *my, *d, *qx, *qy, *om, *qx1, *qy1, *res, *res_next, *src;
};

template <class R>
__host__ __device__ __forceinline__ Cur<R> view(const Bufs<R>& b, int par) {
    // selects between both pointers, not an index into the argument block: the pointers load at the kernel's start
    // instead of after the parity
    const bool p = (par & 1) != 0;
    Cur<R> c;
    c.h = p ? b.h[1] : b.h[0];
    c.h_next = p ? b.h[0] : b.h[1];
    c.u = p ? b.u[1] : b.u[0];
    c.u1 = p ? b.u[0] : b.u[1];
    c.v = p ? b.v[1] : b.v[0];
    c.v1 = p ? b.v[0] : b.v[1];
    c.mx = p ? b.mx[1] : b.mx[0];
    c.my = p ? b.my[1] : b.my[0];
    c.d = b.d;
    c.qx = b.qx;
    c.qy = b.qy;
    c.om = b.om;
    c.qx1 = p ? b.qx1[1] : b.qx1[0];  // written by this step's faces, read by its update
    c.qy1 = p ? b.qy1[1] : b.qy1[0];
    c.res = p ? b.res[1] : b.res[0];
    c.res_next = p ? b.res[0] : b.res[1];
    c.src = b.src;
    return c;
}

struct Statics {
    const float* hc;          // ncell
    const float2* ctab;       // (rows - 1) planes of ncell
    const float4* ftabx;      // (rows - 1) planes of ncell
    const float4* ftaby;
    const float4* sx;         // (fl, fr, gl, gr) of x faces
    const float4* sy;
    const float2* es;         // (efx, efy) table tops, negative for a face with a crest coefficient
    const float2* cwv;        // (cwx, cwy), read only for such a face
    const float* zcx;         // ncell, own layout; a neighbour tile's value is read from that tile
    const float* zcy;
    float4* fcx;              // row caches of the face tables and their row tags (255: empty)
    float4* fcy;
    unsigned char* ftagx;
    unsigned char* ftagy;
    float2* cc;               // row cache of the cell tables
    unsigned char* ctag;
    const int* nbr;           // (ntiles, 8)
    size_t ncell;
    int last;                 // rows - 1
};

// Workgroups go round-robin over the 8 XCDs, each with its own L2. Block b takes slot k = b / 8 of its XCD's
// contiguous share of the n listed tiles, so a tile's east and west neighbours (the next and previous entries of the
// list) are read by the same L2 at about the same time; -1 for a block with no slot.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_slot(int b, int n) {
    const int x = b % XCDS, k = b / XCDS, q = n / XCDS, r = n % XCDS;
    if (k >= q + (x < r ? 1 : 0)) return -1;
    return x * q + (x < r ? x : r) + k;
}

// The 3 x 3 block of tiles around t, (dj + 1) * 3 + (di + 1); NEIGHBOURS order of grid/tiles.py:12.
__device__ __forceinline__ void load_nb3(int t, const int* __restrict__ nbr, int* nb) {
    if (threadIdx.x < 9) {
        const int q = threadIdx.x;
        nb[q] = q == 4 ? t : nbr[(size_t)t * 8 + (q < 4 ? q : q - 1)];
    }
}

// LDS block of one array: 36 rows of LS entries; the tile's own cells at rows 2..33, columns 4..35 (16-byte
// aligned rows), the 2-cell ring at rows 0,1 and 34,35 and columns 2,3 and 36,37.
constexpr int LS = 40;
constexpr int LROWS = TILE + 2 * HALO;
constexpr int LSIZE = LROWS * LS;
constexpr int NRING = STRIDE2 - TILE2;  // 272 ring entries

__device__ __forceinline__ int lds_of(int j, int i) { return (j + HALO) * LS + i + 4; }

// Ring entry q (0 .. 271) as (j, i) of the tile: north rows, south rows, west columns, east columns.
__device__ __forceinline__ void ring_cell(int q, int& j, int& i) {
    if (q < 72) {
        j = q / 36 - 2;
        i = q - (q / 36) * 36 - 2;
    } else if (q < 144) {
        q -= 72;
        j = TILE + q / 36;
        i = q - (q / 36) * 36 - 2;
    } else if (q < 208) {
        q -= 144;
        j = q >> 1;
        i = (q & 1) - 2;
    } else {
        q -= 208;
        j = q >> 1;
        i = TILE + (q & 1);
    }
}

// What one thread loads into every staged array: OWN = 1024 / NTH consecutive own cells and up to RING ring
// entries; the same offsets serve every array, so they are worked out once per workgroup.
template <int NTH>
struct Stager {
    static constexpr int OWN = TILE2 / NTH;
    static constexpr int RING = (NRING + NTH - 1) / NTH;
    int own_g, own_l;
    int ring_g[RING], ring_l[RING];
};

template <int NTH>
__device__ __forceinline__ Stager<NTH> make_stager(int t, const int* nb) {
    Stager<NTH> s;
    const int q0 = threadIdx.x;
    const int p = Stager<NTH>::OWN * q0;
    s.own_g = t * TILE2 + p;
    s.own_l = lds_of(p >> 5, p & (TILE - 1));
#pragma unroll
    for (int m = 0; m < Stager<NTH>::RING; ++m) {
        const int q = q0 + m * NTH;
        s.ring_l[m] = -1;
        s.ring_g[m] = 0;
        if (q < NRING) {
            int j, i;
            ring_cell(q, j, i);
            const int dj = j < 0 ? -1 : (j >= TILE ? 1 : 0);
            const int di = i < 0 ? -1 : (i >= TILE ? 1 : 0);
            const int tt = nb[(dj + 1) * 3 + di + 1];
            s.ring_g[m] = tt * TILE2 + (j - dj * TILE) * TILE + (i - di * TILE);
            s.ring_l[m] = lds_of(j, i);
        }
    }
    return s;
}

// W consecutive values, global -> registers -> LDS, as 16-byte (or 8-byte for two floats) accesses; one value for a
// workgroup of 1024 threads.
template <int W>
__device__ __forceinline__ void ldw(const float* __restrict__ g, float* o) {
    if constexpr (W == 4) {
        const float4 v = *reinterpret_cast<const float4*>(g);
        o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w;
    } else if constexpr (W == 2) {
        const float2 v = *reinterpret_cast<const float2*>(g);
        o[0] = v.x, o[1] = v.y;
    } else {
        static_assert(W == 1);
        o[0] = *g;
    }
}
template <int W>
__device__ __forceinline__ void ldw(const double* __restrict__ g, double* o) {
    if constexpr (W == 1) {
        o[0] = *g;
    } else {
#pragma unroll
        for (int k = 0; k < W; k += 2) {
            const double2 a = *reinterpret_cast<const double2*>(g + k);
            o[k] = a.x, o[k + 1] = a.y;
        }
    }
}
template <int W>
__device__ __forceinline__ void stw(float* l, const float* o) {
    if constexpr (W == 4)
        *reinterpret_cast<float4*>(l) = make_float4(o[0], o[1], o[2], o[3]);
    else if constexpr (W == 2)
        *reinterpret_cast<float2*>(l) = make_float2(o[0], o[1]);
    else
        *l = o[0];
}
template <int W>
__device__ __forceinline__ void stw(double* l, const double* o) {
    if constexpr (W == 1) {
        *l = o[0];
    } else {
#pragma unroll
        for (int k = 0; k < W; k += 2) *reinterpret_cast<double2*>(l + k) = make_double2(o[k], o[k + 1]);
    }
}

// N arrays into their LDS blocks: every load issued before the first store.
template <class R, int N, int NTH>
__device__ __forceinline__ void stage(R* const (&lds)[N], const R* const (&g)[N], const Stager<NTH>& s) {
    constexpr int W = Stager<NTH>::OWN, RN = Stager<NTH>::RING;
    R own[N][W], ring[N][RN];
#pragma unroll
    for (int a = 0; a < N; ++a) {
        ldw<W>(g[a] + s.own_g, own[a]);
#pragma unroll
        for (int m = 0; m < RN; ++m) ring[a][m] = s.ring_l[m] >= 0 ? g[a][s.ring_g[m]] : R(0);
    }
#pragma unroll
    for (int a = 0; a < N; ++a) {
        stw<W>(lds[a] + s.own_l, own[a]);
#pragma unroll
        for (int m = 0; m < RN; ++m)
            if (s.ring_l[m] >= 0) lds[a][s.ring_l[m]] = ring[a][m];
    }
}

// State index of the west / north neighbour of own cell (j, i) of tile t.
__device__ __forceinline__ unsigned west_of(int t, int j, int i, const int* nb) {
    return i > 0 ? t * TILE2 + j * TILE + i - 1 : nb[3] * TILE2 + j * TILE + (TILE - 1);
}
__device__ __forceinline__ unsigned north_of(int t, int j, int i, const int* nb) {
    return j > 0 ? t * TILE2 + (j - 1) * TILE + i : nb[1] * TILE2 + (TILE - 1) * TILE + i;
}

template <class R>
struct FaceIn {
    const R* su;   // normal velocity (LDS)
    const R* sv;   // transverse velocity (LDS)
    const R* sq;   // normal discharge (LDS)
    const R* sqp;  // transverse discharge (LDS)
    const R* sh;
    const R* sd;
    int S, P;      // LDS steps along and across the normal
};

// One face: the west (x) or north (y) face of own cell L of the LDS block; faces.cu:105-242 restated, with the
// statement order of solver/cpu/kernels.py:86-219. Returns the wave speed. LEARNED: the learned closure's
// correction of this face, `corr` = (conveyance factor on T, rise of the face's inverts), is applied (no counterpart
// in the reference; the plain path never takes this branch).
// `zc_up`, `zc_own`: the centre-line beds of the face's two cells (west or north, own); `om_up`, `om_own`: their
// wet plan shares, read only on a wave step. The caller loads them before the face's upwind side is known, so no
// load waits on it.
template <class R, bool FAST, bool LEARNED = false>
__device__ __forceinline__ R face(const FaceIn<R>& in, int L, R m_own, float4 st, R ef, R cw, R zc_up, R zc_own,
                                  const float4* __restrict__ ftab, size_t plane, int last, float4* cache,
                                  unsigned char* tag, R om_up, R om_own, R dt, bool want_wave, const Params& P,
                                  R& u_out, R& m_out, R& q_out, float2 corr = float2{1.f, 0.f}) {
    const int S = in.S, Q = in.P;
    const R* u = in.su;
    const R g = R(P.g), dx = R(P.dx);
    R fl = R(st.x), fr = R(st.y), gl = R(st.z), gr = R(st.w);
    if constexpr (LEARNED) {
        fl += R(corr.y);
        fr += R(corr.y);
        gl += R(corr.y);
        gr += R(corr.y);
    }
    const R eL = in.sd[L - S] - fl;
    const R eR = in.sd[L] - fr;
    const R uf = u[L];
    // the face's row cache, read with the statics: the row is known only after the carry
    const int have = cache ? (int)ld_nt(tag) : -1;
    const float4 pre = cache ? ld_nt(cache) : float4{0.f, 0.f, 0.f, 0.f};
    const R eLp = fmx<FAST>(eL, R(0));
    const R eRp = fmx<FAST>(eR, R(0));
    const bool from_left = uf > R(0) || (uf == R(0) && eL >= eR);
    const R eup = from_left ? eLp : eRp;
    if (eup <= R(P.e_wet)) {
        u_out = R(0);
        m_out = R(0);
        q_out = R(0);
        return R(0);
    }
    R eface = eup;
    if (P.carry && cw == R(0)) {
        const R dz = from_left ? zc_up - fl : zc_own - fr;
        if (dz > R(1e-6)) {
            R phi = dv<FAST>(eup - R(2) * dz, R(2) * dz);
            phi = phi < R(0) ? R(0) : (phi > R(1) ? R(1) : phi);
            eface = eup - phi * dz;
        }
    }
    R A, T, W;
    face_lookup<R, FAST>(eface, ef, ftab, plane, last, A, T, W, cache, tag, have, pre);
    if constexpr (LEARNED) T *= R(corr.x);
    const R hL = in.sh[L - S];
    const R hR = in.sh[L];
    const R hup = from_left ? hL : hR;

    // [1] section
    R rho = R(1);
    R dEta = eRp - eLp;
    if (P.sheet) {
        const R Avol = hup * dx;
        const R edown = from_left ? eRp : eLp;
        if (Avol < A) {
            const R s = dv<FAST>(fmn<FAST>(eLp, eRp), fmx<FAST>(eLp, eRp));
            const R s2 = s * s;
            const R alpha = s2 * s2;
            rho = alpha + dv<FAST>((R(1) - alpha) * Avol, A);
        }
        if (edown <= R(0)) {
            // [2] disconnected
            const R gU = from_left ? gl : gr;
            const R gD = from_left ? gr : gl;
            const R hD = from_left ? hR : hL;
            const R drop = (hup - gU) - fmx<FAST>(hD - gD, R(0.5) * (gU - gD));
            const R level_drop = fabs(eL - eR);
            const R drive = fmx<FAST>(eup, fmn<FAST>(drop, level_drop));
            dEta = from_left ? -drive : drive;
        }
    }

    // [3] upwinded momentum fluxes, along then across
    const bool mirror = P.mirror, second = P.second;
    const R um = gated(u[L - S], uf, mirror);
    const R up = gated(u[L + S], uf, mirror);
    const R umm = gated(u[L - 2 * S], um, mirror);
    const R upp = gated(u[L + 2 * S], up, mirror);
    const R* q = in.sq;
    const R QcL = R(0.5) * (q[L - S] + q[L]);
    const R QcR = R(0.5) * (q[L] + q[L + S]);
    // each centre velocity: the upwind face's, corrected; the arguments are chosen first, one correction is made
    const bool pL = QcL > R(0), pR = QcR > R(0);
    const R sL = ustar(pL ? um : uf, pL ? umm : up, pL ? uf : um, second);
    const R sR = ustar(pR ? uf : up, pR ? um : upp, pR ? up : uf, second);
    R flux = QcR * sR - QcL * sL;
    const R* qp = in.sqp;
    const R QyS = R(0.5) * (qp[L - S] + qp[L]);
    const R QyN = R(0.5) * (qp[L - S + Q] + qp[L + Q]);
    const R uS = gated(u[L - Q], uf, mirror);
    const R uN = gated(u[L + Q], uf, mirror);
    const R uSS = gated(u[L - 2 * Q], uS, mirror);
    const R uNN = gated(u[L + 2 * Q], uN, mirror);
    const bool pS = QyS > R(0), pN = QyN > R(0);
    const R usS = ustar(pS ? uS : uf, pS ? uSS : uN, pS ? uf : uS, second);
    const R usN = ustar(pN ? uf : uN, pN ? uS : uNN, pN ? uN : uf, second);
    flux += QyN * usN - QyS * usS;

    // [4] momentum and pressure
    const R Abar = R(0.5) * (hL + hR) * dx;
    const R press = dv<FAST>(dt * g * dEta, dx);
    R us;
    if (Abar > R(1e-6) * dx)
        us = dv<FAST>(m_own - dv<FAST>(dt * flux, dx), Abar) - press;
    else
        us = uf - press;

    // [5] friction, implicit in the speed
    const R* vp = in.sv;
    const R vb = R(0.25) * (vp[L - S] + vp[L] + vp[L - S + Q] + vp[L + Q]);
    const R T3 = T * T * T * cb<FAST>(T);
    R fric = dv<FAST>(A * A, fmx<FAST>(T3, R(1e-30)));
    if (rho < R(1)) fric = dv<FAST>(fric, fmx<FAST>(rho * cb<FAST>(rho), R(1e-12)));
    const R a = dt * g * fric;
    const R sp = sq<FAST>(us * us + vb * vb);
    R un = dv<FAST>(R(2) * us, R(1) + sq<FAST>(R(1) + R(4) * a * sp));
    const R umax = R(P.u_max);
    if (fabs(un) > umax) un = un > R(0) ? umax : -umax;

    // [6] crest face: no more than critical flow over the crest
    R qn = A * rho * un;
    R c = cw;
    if (c != R(0)) {
        R Ac = A, Wc = W;
        if (c < R(0)) {
            c = -c;
            const R ua = from_left ? um : up;
            R Tc;
            face_lookup<R, FAST>(eup + dv<FAST>(R(0.5) * ua * ua, g), ef, ftab, plane, last, Ac, Tc, Wc);
        }
        const R qc = c * Ac * sq<FAST>(dv<FAST>(Ac, fmx<FAST>(Wc, R(1e-9))));
        if (fabs(qn) > qc) {
            qn = qn > R(0) ? qc : -qc;
            un = dv<FAST>(qn, fmx<FAST>(A * rho, R(1e-12)));
        }
    }
    u_out = un;
    m_out = Abar * un;
    q_out = qn;
    if (!want_wave) return R(0);
    if (!isfinite(un)) return R(1e30);
    // wave speed for the next refresh
    const R dL = in.sd[L - S];
    const R dR = in.sd[L];
    const R dup = from_left ? dL : dR;
    const R h_cfl = R(P.h_cfl);
    if (dup <= h_cfl) return R(0);
    R plan = from_left ? om_up : om_own;
    const R dother = from_left ? dR : dL;
    if (dother > h_cfl) plan = fmn<FAST>(plan, from_left ? om_own : om_up);
    const R wave = sq<FAST>(fmx<FAST>(dv<FAST>(g * A * rho, fmx<FAST>(plan, R(P.a_min)) * dx), R(0)));
    return fabs(un) + fmn<FAST>(R(P.lead) * fabs(un - uf), R(2) * wave) + wave;
}

template <int NTH, class R>
__device__ __forceinline__ R block_max(R v, R* red) {
    for (int off = 32; off > 0; off >>= 1) {
        const R o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wid] = v;
    __syncthreads();
    R m = red[0];
    for (int w = 1; w < NTH / 64; ++w) m = red[w] > m ? red[w] : m;
    return m;
}

// A wave speed (never negative, never NaN: the maxima above skip NaN) as bits whose unsigned order is its order, for
// an atomic max across workgroups; and back.
__device__ __forceinline__ unsigned long long wave_bits(float v) { return (unsigned long long)__float_as_uint(v); }
__device__ __forceinline__ unsigned long long wave_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
template <class R>
__device__ __forceinline__ double wave_value(unsigned long long b) {
    if constexpr (std::is_same<R, float>::value) return (double)__uint_as_float((unsigned)b);
    else return __longlong_as_double((long long)b);
}

// One member of a device-resident run: a model with its own clock, step, source windows and books. A plain run
// is one member; an ensemble is many, their tiles side by side in one set of arrays (tile0 .. tile1).
struct Member {
    double t, t_end, dt, dt_step;
    double window_start, window_end, window_len, scale;
    double rain_done, rain_window;  // FP64 rain volume: whole past windows, and the current window
    long long nsteps;
    long long scale_off;
    double dt_min;  // the smallest dt held after a step of this run
    double t_stop;  // the member holds (steps no further) once its clock has reached this; the host moves it on
    int nwin, win;
    int tile0, tile1;
    int b0, b1, f0, f1;  // the member's boundary cells and tally faces in the handle's lists
    int stepping, refresh, refresh_due, new_window, dry, rebuild, check, done, failed;
    int nlist;  // the member's active tiles, kept at seg[tile0 .. tile0 + nlist)
    int par;  // the member's buffer parity (see Cur): a member that does not step keeps its own
    double q_peak, t_peak;  // the largest discharge out through the member's tally faces after a step, and when
    // The pipelined step (pipe_kernel): with `pipe`, a step whose next step keeps the list and reads no wet flag
    // (`fuse`) leaves its cell update to the next step's pipe kernel; `pend` says an update waits there, with its
    // step's parameters (`u_*`: length, wave step, new window and scale, parity, steps before it, clock after it).
    int pipe, fuse, pend;
    int kept_rain, kept_dry;  // how many of the member's tiles a check keeps whatever their wet flags say: while it
                              // rains (a source or a boundary cell in the tile) and when dry (a boundary cell)
    int u_want, u_new_window, u_par;
    long long u_nsteps;
    long long u_made;  // updates the pipe kernel has made for this member (a count for the host)
    double u_dt, u_scale, u_time;
};

struct RunGlobals {
    int nact;
    int changed;   // a member's list changed since the active list was last gathered
    int pre_done;  // workgroups of the pre kernel that have finished (the last gathers the lists, then clears it)
    int classic;   // a member's update this step is the cell kernel's (set by the pre kernel, cleared by the post kernel)
};

struct RunNum {
    double cfl, dt_max, dt_min, wave_floor, dx, area;
    int dt_every, active_every;
    int excess;  // the source is the rain excess of each cell (excess_kernel), not weight * scale
};

template <class R>
struct StepArgs {
    Bufs<R> b;
    int par;  // a protocol step's buffer parity (a device-resident run takes each member's)
    Statics st;
    Params P;
    const int* active;
    int nact;
    R dt;
    int want_wave;
    R* wave;     // per tile
    int* wet;    // per tile
    // device-resident run (null in a protocol step): each tile's member says its step, flags and source
    const Member* members;
    const int* tile_member;
    const RunGlobals* glob;
    const double* src_weight;
    double* src_partial;
    int excess;  // device-resident run with the excess source: excess_kernel has written the window's rates
    // the learned closure's per-face corrections (null when it is off)
    const float2* corr_x;
    const float2* corr_y;
    // device-resident run: each member's largest wave speed of the step, as wave_bits (null in a protocol step)
    unsigned long long* member_wave;
    // face momentum: fresh says mx, my hold it (set from the host); else it is the last step's Abar * u, Abar from
    // the mean depths that step read (h_old), recomputed here (a step never stores it)
    int fresh;
    // device-resident run with the running maxima (every > 0): the cell kernel samples them after the steps whose
    // count is a multiple of `every`, on every tile without a boundary cell (bnd_tile); those tiles are sampled after
    // the boundary cells are set (maxima_run_kernel)
    int every;
    R arrive, deep;
    R *mx_d, *mx_h, *mx_s, *mx_z, *mx_t;
    const int* bnd_tile;
};

// Both faces of own cell c (LDS entry L) of a tile: the x (west) face, then the y (north) face (faces.cu:244-271),
// from the cell's momenta and the beds and wet shares of the face's two cells (west or north, own); writes the next
// face velocities and the discharges before the limit of the view `s`, returns the larger wave speed.
template <class R, bool FAST, bool LEARNED>
__device__ __forceinline__ R face_cell(const StepArgs<R>& A, const Cur<R>& s, const FaceIn<R>& fx,
                                       const FaceIn<R>& fy, int L, unsigned c, R mxo, R myo, R zxw, R zxo, R zyn,
                                       R zyo, R omw, R omn, R omo, R dt, bool want_wave) {
    const float2 e = ld_nt(A.st.es + c);
    const R efx = R(fabsf(e.x)), efy = R(fabsf(e.y));
    const R cwx = signbit(e.x) ? R(A.st.cwv[c].x) : R(0);
    const R cwy = signbit(e.y) ? R(A.st.cwv[c].y) : R(0);
    R un, mn, qn;
    float2 cx{1.f, 0.f}, cy{1.f, 0.f};
    if constexpr (LEARNED) {
        cx = A.corr_x[c];
        cy = A.corr_y[c];
    }
    const R wx = face<R, FAST, LEARNED>(fx, L, mxo, ld_nt(A.st.sx + c), efx, cwx, zxw, zxo, A.st.ftabx + c,
                                        A.st.ncell, A.st.last, A.st.fcx ? A.st.fcx + c : nullptr, A.st.ftagx + c, omw,
                                        omo, dt, want_wave, A.P, un, mn, qn, cx);
    st_nt(un, s.u1 + c);
    st_nt(qn, s.qx1 + c);
    const R wy = face<R, FAST, LEARNED>(fy, L, myo, ld_nt(A.st.sy + c), efy, cwy, zyn, zyo, A.st.ftaby + c,
                                        A.st.ncell, A.st.last, A.st.fcy ? A.st.fcy + c : nullptr, A.st.ftagy + c, omn,
                                        omo, dt, want_wave, A.P, un, mn, qn, cy);
    st_nt(un, s.v1 + c);
    st_nt(qn, s.qy1 + c);
    return wx > wy ? wx : wy;
}

// The tile's largest wave speed, kept per tile and raised into its member's (a wave step).
template <class R, int NTH>
__device__ __forceinline__ void tile_wave(const StepArgs<R>& A, int t, int member, R w, R* red) {
    const R m = block_max<NTH>(w, red);
    if (threadIdx.x == 0) {
        A.wave[t] = m;
        if (A.member_wave) atomicMax(A.member_wave + member, wave_bits(m));
    }
}

// The face step of tile t at parity `par` on the stored state: its arrays with the 2-cell ring into LDS (nb: the 3 x 3
// block of tiles around t, loaded), then both own faces of every own cell.
template <class R, bool FAST, int NTH, bool LEARNED>
__device__ __forceinline__ void faces_tile(const StepArgs<R>& A, int t, int member, int par, R dt, bool want_wave,
                                           const int* nb, R* s_u, R* s_v, R* s_h, R* s_d, R* s_qx, R* s_qy, R* red) {
    const Cur<R> s = view(A.b, par);
    __syncthreads();
    {
        const Stager<NTH> sg = make_stager<NTH>(t, nb);
        R* const dst[6] = {s_u, s_v, s_h, s_d, s_qx, s_qy};
        const R* const src[6] = {s.u, s.v, s.h, s.d, s.qx, s.qy};
        stage<R, 6, NTH>(dst, src, sg);
    }
    __syncthreads();
    const FaceIn<R> fx{s_u, s_v, s_qx, s_qy, s_h, s_d, 1, LS};
    const FaceIn<R> fy{s_v, s_u, s_qy, s_qx, s_h, s_d, LS, 1};
    const int i = threadIdx.x & (TILE - 1);
    const bool fresh = A.fresh != 0;
    const R dxr = R(A.P.dx);
    R w = R(0);
#pragma unroll
    for (int r = 0; r < TILE2 / NTH; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const int L = lds_of(j, i);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        // the face momenta the last step left: Abar * u with Abar = 0.5 (hL + hR) dx from the depths it read, the
        // same operations on the same operands as its own product, so the same bits as a stored copy
        R mxo, myo;
        if (fresh) {
            mxo = s.mx[c];
            myo = s.my[c];
        } else {
            const R ho = ld_nt(s.h_next + c);
            mxo = R(0.5) * (s.h_next[west_of(t, j, i, nb)] + ho) * dxr * s_u[L];
            myo = R(0.5) * (s.h_next[north_of(t, j, i, nb)] + ho) * dxr * s_v[L];
        }
        const unsigned kw = west_of(t, j, i, nb), kn = north_of(t, j, i, nb);
        const R zxw = R(A.st.zcx[kw]), zxo = R(A.st.zcx[c]), zyn = R(A.st.zcy[kn]), zyo = R(A.st.zcy[c]);
        R omw = R(0), omn = R(0), omo = R(0);
        if (want_wave) {
            omw = s.om[kw];
            omn = s.om[kn];
            omo = s.om[c];
        }
        const R m = face_cell<R, FAST, LEARNED>(A, s, fx, fy, L, c, mxo, myo, zxw, zxo, zyn, zyo, omw, omn, omo, dt,
                                                want_wave);
        w = m > w ? m : w;
    }
    if (want_wave) tile_wave<R, NTH>(A, t, member, w, red);
}

// Both own faces of every own cell of one active tile per workgroup of NTH threads (faces.cu:244-271).
template <class R, bool FAST, int NTH, bool LEARNED = false>
__global__ __launch_bounds__(NTH) void faces_kernel(StepArgs<R> A) {
    __shared__ __align__(16) R s_u[LSIZE];
    __shared__ __align__(16) R s_v[LSIZE];
    __shared__ __align__(16) R s_h[LSIZE];
    __shared__ __align__(16) R s_d[LSIZE];
    __shared__ __align__(16) R s_qx[LSIZE];
    __shared__ __align__(16) R s_qy[LSIZE];
    __shared__ int nb[9];
    __shared__ R red[NTH / 64];
    const int slot = xcd_slot(blockIdx.x, A.members ? A.glob->nact : A.nact);
    if (slot < 0) return;
    const int t = uni(A.active[slot]);
    R dt_run = A.dt;
    int want_run = A.want_wave;
    int member = 0, par = A.par;
    if (A.members) {
        member = uni(A.tile_member[t]);
        const Member& mb = A.members[member];
        if (!uni(mb.stepping)) return;
        dt_run = (R)mb.dt_step;
        want_run = uni(mb.refresh);
        par = uni(mb.par);
    }
    load_nb3(t, A.st.nbr, nb);
    faces_tile<R, FAST, NTH, LEARNED>(A, t, member, par, dt_run, want_run != 0, nb, s_u, s_v, s_h, s_d, s_qx, s_qy,
                                      red);
}

// The learned closure on its faces, before a step that refreshes the wave speed (every dt_every steps and after a
// source change; levels move little between): each wave takes 16 faces, forms
// their 16 features (11 static, then e_L, e_R, h_L, h_R and the upwind level over the crest, from the current level
// and depth), standardises them, evaluates the network on the matrix cores and leaves the face's correction:
// conveyance factor exp(clamp(y0)) (1 unless `conveyance`) and invert rise clamp(y1 * scale, 0, 4 m).
struct ClosureArgs {
    int n;
    const int* cell_r;  // the face's own cell (whose west or north face it is), own layout
    const int* cell_l;  // the cell on its other side
    const int* axis;    // 1: x face, 0: y face
    const float* stat;  // n x 11
    const float* norm;  // 16 means, 16 inverse deviations, then the scale on the invert-rise output
    const uint16_t* wbf;
    const float* params;
    const float4* sx;
    const float4* sy;
    float2* corr_x;
    float2* corr_y;
    int conveyance;
    // device-resident run: only the faces of members on a refresh step are evaluated (null: every face)
    const Member* members;
    const int* tile_member;
};

constexpr int CLOSURE_STATIC = 11;
constexpr float CLOSURE_LOGC = 1.5f, CLOSURE_RISE = 4.f;

__global__ __launch_bounds__(256) void closure_kernel(ClosureArgs C, Bufs<float> b, int par) {
    using namespace w2d_mlp;
    __shared__ Net net;
    __shared__ __align__(16) uint16_t xs[4][WAVE_SAMPLES * IN];
    __shared__ __align__(16) uint16_t h1[4][WAVE_SAMPLES * HID];
    __shared__ __align__(16) uint16_t h2[4][WAVE_SAMPLES * HID];
    __shared__ float ys[4][WAVE_SAMPLES * OUT];
    load_net(net, C.wbf, C.params);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int base = blockIdx.x * 64; base < C.n; base += gridDim.x * 64) {
        const int f = base + wv * WAVE_SAMPLES + (lane & 15);
        if (C.members) {  // a wave whose faces all belong to members not refreshing this step has nothing to do
            bool due = false;
            if (lane < 16 && f < C.n) {
                const Member& mb = C.members[C.tile_member[C.cell_r[f] / TILE2]];
                due = mb.stepping && mb.refresh;
            }
            if (!__any(due)) continue;
        }
        if (lane < 16) {
            float x[IN];
            if (f < C.n) {
                const int r = C.cell_r[f], l = C.cell_l[f];
                // the current level and depth of the face's member (a protocol step: the handle's parity)
                const int pr = C.members ? C.members[C.tile_member[r / TILE2]].par : par;
                const float* __restrict__ d = b.d;
                const float* __restrict__ h = b.h[pr & 1];
                const float4 st = C.axis[f] ? C.sx[r] : C.sy[r];
                const float* s = C.stat + (size_t)f * CLOSURE_STATIC;
                for (int q = 0; q < CLOSURE_STATIC; ++q) x[q] = s[q];
                const float eL = fmaxf(d[l] - st.x, 0.f), eR = fmaxf(d[r] - st.y, 0.f);
                x[11] = eL;
                x[12] = eR;
                x[13] = h[l];
                x[14] = h[r];
                x[15] = fmaxf(eL, eR) - s[5];
            } else {
                for (int q = 0; q < IN; ++q) x[q] = C.norm[q];
            }
            for (int q = 0; q < IN; ++q) xs[wv][lane * IN + q] = to_bf16((x[q] - C.norm[q]) * C.norm[IN + q]);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        forward_wave(net, xs[wv], h1[wv], h2[wv], ys[wv]);
        if (lane < 16 && f < C.n) {
            const float y0 = ys[wv][lane * OUT], y1 = ys[wv][lane * OUT + 1];
            const float c = C.conveyance ? expf(fminf(fmaxf(y0, -CLOSURE_LOGC), CLOSURE_LOGC)) : 1.f;
            const float rise = fminf(fmaxf(y1 * C.norm[2 * IN], 0.f), CLOSURE_RISE);
            (C.axis[f] ? C.corr_x : C.corr_y)[C.cell_r[f]] = make_float2(c, rise);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
}

template <class R>
__device__ __forceinline__ R scaled(R q, R rl, R rr) { return q * (q > R(0) ? rl : rr); }

// LDS index of tile cell (j, i) in a block whose row 0 holds tile row -ROW0 (LS entries per row, own cells from
// column 4): ROW0 = 2 for the two-pass kernels' blocks (lds_of), 3 for the fused step's.
template <int ROW0>
__device__ __forceinline__ int lidx(int j, int i) { return (j + ROW0) * LS + i + 4; }

// The positivity ratio (cells.cu:4-18) of the tile and a one-cell ring, into s_r (34 x 34, cell (j, i) at
// (j + 1) * 34 + i + 1), from the mean depth and the face discharges before the limit in LDS. The four corner cells
// are never read; CORNERS says whether their faces are in LDS (the fused step computes no face beyond the ring).
constexpr int RW = TILE + 2;
template <class R, bool FAST, int NTH, int ROW0, bool CORNERS>
__device__ __forceinline__ void limit_ring(const R* s_h, const R* s_qx, const R* s_qy, R* s_r, R dt, R area) {
    for (int p = threadIdx.x; p < RW * RW; p += NTH) {
        const int jj = p / RW, ii = p - (p / RW) * RW;
        if (!CORNERS && (jj == 0 || jj == RW - 1) && (ii == 0 || ii == RW - 1)) continue;
        const int L = lidx<ROW0>(jj - 1, ii - 1);
        const R out = pmax(s_qx[L + 1], R(0)) + pmax(-s_qx[L], R(0)) + pmax(s_qy[L + LS], R(0)) +
                      pmax(-s_qy[L], R(0));
        const R have = pmax(s_h[L], R(0)) * area;
        s_r[p] = out * dt > have ? dv<FAST>(have, out * dt) : R(1);
    }
}

// Continuity, rain, the compensated residual, the level and the wet share of the tile's own cells
// (cells.cu:28-63), from the ratio and the face discharges before the limit in LDS: writes the limited discharges,
// the new mean depth, level and residual, the wet share on a wave step, the window's source; returns whether a cell
// is wet and adds the window's source to src_sum.
// With `sample`, it then takes the running maxima of the tile's cells at time `time` (the statements of
// maxima_run_kernel, on the values just computed; `nb` the 3 x 3 block of tiles around t).
template <class R, bool FAST, int NTH, int ROW0>
__device__ __forceinline__ bool update_tile(const StepArgs<R>& A, const Cur<R>& s, int t, const R* s_h,
                                            const R* s_qx, const R* s_qy, const R* s_r, R dt, bool want_plan,
                                            bool new_window, double scale, double& src_sum, bool sample = false,
                                            R time = R(0), const int* nb = nullptr) {
    const R dx = R(A.P.dx);
    const R area = dx * dx;
    const int i = threadIdx.x & (TILE - 1);
    bool wet = false;
    // The thread's RT cells in three passes, so that the loads of one cell need not wait for the stores of the
    // one before: (1) the cell's inputs, (2) continuity and the level table's row pair, (3) the stores.
    constexpr int RT = TILE2 / NTH;
    unsigned cs[RT];
    R res_v[RT], src_v[RT];
    float hc_v[RT];
    int tag_v[RT];
    R md_v[RT], mh_v[RT], mt_v[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        cs[r] = c;
        res_v[r] = ld_nt(s.res + c);
        hc_v[r] = ld_nt(A.st.hc + c);
        tag_v[r] = ld_nt(A.st.ctag + c);
        if (new_window)  // the window's rate of this cell: float(weight * scale), the product in double
            src_v[r] = A.src_weight ? (R)(A.src_weight[c] * scale) : R(0);
        else
            src_v[r] = ld_nt(s.src + c);
        if (sample) {
            md_v[r] = A.mx_d[c];
            mh_v[r] = A.mx_h[c];
            mt_v[r] = A.mx_t[c];
        }
    }
    R qw_v[RT], qn_v[RT], hn_v[RT], resn_v[RT], s_v[RT], dn_v[RT];
    int row_v[RT], mode_v[RT];
    float2 pair_v[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const int L = lidx<ROW0>(j, i);
        const int Q = (j + 1) * RW + (i + 1);  // the cell in s_r
        const unsigned c = cs[r];
        const R rc = s_r[Q];
        const R qw = scaled(s_qx[L], s_r[Q - 1], rc);
        const R qe = scaled(s_qx[L + 1], rc, s_r[Q + 1]);
        const R qn = scaled(s_qy[L], s_r[Q - RW], rc);
        const R qs = scaled(s_qy[L + LS], rc, s_r[Q + RW]);
        qw_v[r] = qw;
        qn_v[r] = qn;
        const R h0 = s_h[L];
        const R inc = res_v[r] + dt * src_v[r] - dv<FAST>(dt * (qe - qw + qs - qn), area);
        const R sum = h0 + inc;
        const R hn = pmax(sum, R(0));
        resn_v[r] = sum > R(0) ? inc - (sum - h0) : R(0);
        hn_v[r] = hn;
        const R Hc = R(hc_v[r]);
        // the level table's row for hn (the statements of cell_level and cell_plan, the row pair read through the
        // cell's row cache): 0 no table, 1 above its top, 2 within it
        int mode = 0, row = 0;
        R sv = R(0);
        if (Hc > R(0)) {
            if (hn >= Hc) {
                mode = 1;
                row = A.st.last - 1;
            } else {
                mode = 2;
                sv = sq<FAST>(dv<FAST>(hn, Hc)) * R(A.st.last);
                row = row_of(sv, A.st.last - 1);
            }
        }
        mode_v[r] = mode;
        row_v[r] = row;
        s_v[r] = sv;
        if (mode) pair_v[r] = tag_v[r] == row ? A.st.cc[c] : A.st.ctab[(size_t)row * A.st.ncell + c];
    }
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const unsigned c = cs[r];
        st_nt(qw_v[r], s.qx + c);
        st_nt(qn_v[r], s.qy + c);
        if (new_window) {
            s.src[c] = src_v[r];
            src_sum += (double)src_v[r];
        }
        st_nt(resn_v[r], s.res_next + c);
        const R hn = hn_v[r];
        st_nt(hn, s.h_next + c);
        const R Hc = R(hc_v[r]);
        const float2 q = pair_v[r];
        R dn, plan = R(1);
        if (mode_v[r] == 0) {
            dn = hn;
        } else if (mode_v[r] == 1) {
            dn = R(q.y) + (hn - Hc);
        } else {
            const int row = row_v[r];
            if (want_plan) {
                const R m = R(A.st.last);
                plan = dv<FAST>(dv<FAST>(Hc * R(2 * row + 1), m * m), pmax(R(q.y) - R(q.x), R(1e-12)));
            }
            dn = R(q.x) + (s_v[r] - R(row)) * (R(q.y) - R(q.x));
        }
        if (mode_v[r] && tag_v[r] != row_v[r]) {  // refill the cell's row cache
            A.st.cc[c] = q;
            A.st.ctag[c] = (unsigned char)row_v[r];
        }
        st_nt(dn, s.d + c);
        dn_v[r] = dn;
        if (want_plan) st_nt(plan, s.om + c);
        wet |= hn > R(0);
    }
    if (sample) {
        const int te = nb[5], tso = nb[7];  // east and south neighbours
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
            const unsigned c = cs[r];
            const R dk = dn_v[r], hk = hn_v[r];
            if (dk > md_v[r]) A.mx_d[c] = dk;
            if (hk > mh_v[r]) A.mx_h[c] = hk;
            if (dk > A.deep) {
                const unsigned ke = i < TILE - 1 ? c + 1 : (unsigned)(te * TILE2 + j * TILE);
                const unsigned ks = j < TILE - 1 ? c + TILE : (unsigned)(tso * TILE2 + i);
                const R uc = R(0.5) * (s.u1[c] + s.u1[ke]);
                const R vc = R(0.5) * (s.v1[c] + s.v1[ks]);
                const R sp = sqrt(uc * uc + vc * vc);
                if (sp > A.mx_s[c]) A.mx_s[c] = sp;
                if (sp * dk > A.mx_z[c]) A.mx_z[c] = sp * dk;
            }
            if (mt_v[r] < R(0) && dk >= A.arrive) A.mx_t[c] = time;
        }
    }
    return wet;
}

// The tile's wet flag and, on a new window, its source sum (one per tile), after update_tile.
template <int NTH>
__device__ __forceinline__ void tile_flags(int t, bool wet, bool new_window, double src_sum, int* any_wet,
                                           double* red_src, int* wet_out, double* src_partial) {
    if (wet) *any_wet = 1;
    if (new_window) {
        for (int off = 32; off > 0; off >>= 1) src_sum += __shfl_xor(src_sum, off);
        if ((threadIdx.x & 63) == 0) red_src[threadIdx.x >> 6] = src_sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        wet_out[t] = *any_wet;
        if (new_window) {
            double sum = red_src[0];
            for (int w = 1; w < NTH / 64; ++w) sum += red_src[w];
            src_partial[t] = sum;
        }
    }
}

// The run's per-tile step parameters: this step's length, whether it refreshes the wave speed, whether it opens a
// source window (and its scale), the buffer parities; false for a tile whose member does not step.
template <class R>
struct TileStep {
    R dt;
    int want, new_window, member, par;
    long long nsteps;
    double scale;
    int sample;  // the running maxima are sampled after this step, here
    R time;      // the member's time after the step (the post kernel's t + dt_step)
};
template <class R>
__device__ __forceinline__ bool tile_step(const StepArgs<R>& A, int t, TileStep<R>& ts) {
    ts.dt = A.dt;
    ts.want = A.want_wave;
    ts.new_window = 0;
    ts.scale = 0.0;
    ts.member = 0;
    ts.par = A.par;
    ts.nsteps = 0;
    ts.sample = 0;
    ts.time = R(0);
    if (A.members) {
        ts.member = uni(A.tile_member[t]);
        const Member& mb = A.members[ts.member];
        if (!uni(mb.stepping) || uni(mb.fuse)) return false;  // a waiting update is the next pipe kernel's
        ts.dt = (R)mb.dt_step;
        ts.want = uni(mb.refresh);
        ts.new_window = uni(mb.new_window) && !A.excess;
        ts.scale = mb.scale;
        ts.par = uni(mb.par);
        ts.nsteps = mb.nsteps;
        if (A.every > 0 && (mb.nsteps + 1) % A.every == 0 && !A.bnd_tile[t]) {
            ts.sample = 1;
            ts.time = (R)(mb.t + mb.dt_step);
        }
    }
    return true;
}

// Positivity ratio on the tile and a one-cell ring (cells.cu:4-18), then continuity, rain, the compensated
// residual, the level and wet share of every own cell (cells.cu:28-63): the second kernel of the two-pass step.
template <class R, bool FAST, int NTH>
__global__ __launch_bounds__(NTH) void cells_kernel(StepArgs<R> A) {
    if (A.members && !A.glob->classic) return;  // every member's update waits for the next pipe kernel
    __shared__ __align__(16) R s_h[LSIZE];
    __shared__ __align__(16) R s_qx[LSIZE];
    __shared__ __align__(16) R s_qy[LSIZE];
    __shared__ __align__(16) R s_r[RW * RW];  // compact (7 workgroups per CU fit the LDS)
    __shared__ int nb[9];
    __shared__ int any_wet;
    __shared__ double red_src[NTH / 64];
    const int slot = xcd_slot(blockIdx.x, A.members ? A.glob->nact : A.nact);
    if (slot < 0) return;
    const int t = uni(A.active[slot]);
    TileStep<R> ts;
    if (!tile_step(A, t, ts)) return;
    const Cur<R> s = view(A.b, ts.par);
    load_nb3(t, A.st.nbr, nb);
    if (threadIdx.x == 0) any_wet = 0;
    __syncthreads();
    {
        const Stager<NTH> sg = make_stager<NTH>(t, nb);
        R* const dst[3] = {s_h, s_qx, s_qy};
        const R* const src[3] = {s.h, s.qx1, s.qy1};
        stage<R, 3, NTH>(dst, src, sg);
    }
    __syncthreads();
    const R dx = R(A.P.dx);
    limit_ring<R, FAST, NTH, 2, true>(s_h, s_qx, s_qy, s_r, ts.dt, dx * dx);
    __syncthreads();
    double src_sum = 0.0;
    const bool wet = update_tile<R, FAST, NTH, 2>(A, s, t, s_h, s_qx, s_qy, s_r, ts.dt, ts.want != 0,
                                                  ts.new_window != 0, ts.scale, src_sum, ts.sample != 0, ts.time, nb);
    tile_flags<NTH>(t, wet, ts.new_window != 0, src_sum, &any_wet, red_src, A.wet, A.src_partial);
}

// ------------------------------------------------------------------------------------------------------------
// the pipelined step

// What the pipe kernel reads beyond a step's arguments: which tiles are listed (a neighbour tile not listed, or the
// void tile, keeps its stored state), and for the tiles that need them each cell's boundary entry and tally faces.
struct PipeArgs {
    const int* is_active;  // per tile, the device loop's list as flags
    const int* near;       // per tile: bit 0 a boundary cell in the tile or its west or north neighbour, bit 1 a tally
                           // face in the tile
    const int* bidx;       // per cell: its entry in the boundary lists, -1 for none
    const int2* tidx;      // per cell: its x and y tally faces, -1 for none
    const int* b_kind;
    const int* b_ref;
    const float* b_d0;
    const float* b_h0;
    const float* t_sign;
    double* t_total;
    double* t_flow;
    int drain;  // the waiting updates only: no face step
};

// Which of the 3 x 3 tiles around a tile holds its cell (j, i), j and i in [-32, 63], and the cell's state index.
__device__ __forceinline__ int nb_of(int j, int i) {
    return (j < 0 ? 0 : (j >= TILE ? 2 : 1)) * 3 + (i < 0 ? 0 : (i >= TILE ? 2 : 1));
}
__device__ __forceinline__ unsigned cell_of(const int* nb, int j, int i) {
    const int q = nb_of(j, i);
    const int dj = q / 3 - 1, di = q - (q / 3) * 3 - 1;
    return (unsigned)(nb[q] * TILE2 + (j - dj * TILE) * TILE + (i - di * TILE));
}

// One step of a member whose last cell update waits (Member::pend), in one pass over each tile: that update, then this
// step's faces, which read its results from LDS instead of from memory. The update's inputs (the mean depth, residual
// and the discharges before the limit of the step before, with the 2-cell ring) come from the buffers of its own
// parity, which no workgroup writes in this launch; each tile recomputes the update of its west column and north row
// neighbours (whose faces and levels its own faces read) from the same inputs with the same statements as their own
// tile does, and takes the stored state of a neighbour tile that is not listed. A member with nothing waiting takes
// the face step of faces_kernel. With `drain`, only the waiting updates are made.
//   1. positivity ratio on the tile and a 2-cell ring (cells.cu:4-18), from the staged depth and discharges;
//   2. the face momenta of the own faces (Abar * u from the depths the last face step read: the staged ones);
//   3. the discharges limited by the upwind cell's ratio, in place (cells.cu:36-43);
//   4. continuity, rain, residual, level and wet share of the own cells and of the west column and north row
//      (cells.cu:44-63), stored for the own cells, into LDS for all;
//   5. boundary cells among them (cells.cu:95-111; a follower's reference cell is one of them, checked when the run
//      is set up);
//   6. the own cells' running maxima, boundary tally (engine/maxima.py:29-62), wet flag and window source;
//   7. stored discharges of faces whose cell is in a tile that is not listed, as faces_kernel would stage them;
//   8. both faces of every own cell (faces.cu:244-271).
template <class R, bool FAST, int NTH>
__global__ __launch_bounds__(NTH, 6) void pipe_kernel(StepArgs<R> A, PipeArgs G) {
    __shared__ __align__(16) R s_u[LSIZE];
    __shared__ __align__(16) R s_v[LSIZE];
    __shared__ __align__(16) R s_h[LSIZE];  // the old mean depth, then the new one
    __shared__ __align__(16) R s_d[LSIZE];
    __shared__ __align__(16) R s_qx[LSIZE];  // the discharges before the limit, then after it
    __shared__ __align__(16) R s_qy[LSIZE];
    __shared__ __align__(16) R s_r[LSIZE];  // the positivity ratio, then the wet share
    __shared__ R s_mx[TILE2], s_my[TILE2];   // the own faces' momenta (x, y), from the staged depths
    __shared__ int nb[9], live[9];
    __shared__ R red[NTH / 64];
    __shared__ int any_wet;
    __shared__ double red_src[NTH / 64];
    constexpr int RO = TILE2 / NTH;
    const int slot = xcd_slot(blockIdx.x, A.glob->nact);
    if (slot < 0) return;
    const int t = uni(A.active[slot]);
    const int member = uni(A.tile_member[t]);
    const Member& mb = A.members[member];
    const bool pend = uni(mb.pend) != 0;
    const bool step = !G.drain && uni(mb.stepping) != 0;
    if (!pend && !step) return;
    load_nb3(t, A.st.nbr, nb);
    if (!pend) {
        faces_tile<R, FAST, NTH, false>(A, t, member, uni(mb.par), (R)mb.dt_step, uni(mb.refresh) != 0, nb, s_u, s_v,
                                        s_h, s_d, s_qx, s_qy, red);
        return;
    }
    const int up = uni(mb.u_par);
    const Cur<R> su = view(A.b, up);      // the waiting update
    const Cur<R> sf = view(A.b, up ^ 1);  // this step's faces
    const R dt_u = (R)mb.u_dt;
    const bool plan_u = uni(mb.u_want) != 0;
    const bool win_u = uni(mb.u_new_window) != 0 && !A.excess;
    const double scale_u = mb.u_scale;
    const bool sample = A.every > 0 && (mb.u_nsteps + 1) % A.every == 0;
    const bool want_wave = step && uni(mb.refresh) != 0;
    const int flags = uni(G.near[t]);
    const bool failed = uni(mb.failed) != 0;  // the waiting step failed: its tally is not booked
    // the level and limited discharges of the update are stored only when no face step follows in this launch (a
    // drain, a held or failed member): a following step's update, pipelined or not, rewrites them before anything
    // reads them, and a tile leaves the list only after a step whose update the cell kernel made
    const bool keep = !step;
    if (threadIdx.x < 9) live[threadIdx.x] = nb[threadIdx.x] != 0 && G.is_active[nb[threadIdx.x]] != 0;
    if (threadIdx.x == 0) any_wet = 0;
    __syncthreads();
    {
        const Stager<NTH> sg = make_stager<NTH>(t, nb);
        R* const dst[5] = {s_h, s_qx, s_qy, s_u, s_v};
        const R* const src[5] = {su.h, su.qx1, su.qy1, sf.u, sf.v};
        stage<R, 5, NTH>(dst, src, sg);
    }
    __syncthreads();
    const R dx = R(A.P.dx);
    const R area = dx * dx;
    const int i = threadIdx.x & (TILE - 1);
    // the update's inputs from memory, loaded now so that they arrive during phases 1-3: the own cells' residual,
    // source and table top (ordinary loads: the east and south neighbours read the edge ones again) and row tag ...
    unsigned cs[RO];
    R res_v[RO], src_v[RO];
    float hc_v[RO];
    int tag_v[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        cs[r] = c;
        res_v[r] = su.res[c];
        hc_v[r] = A.st.hc[c];
        tag_v[r] = ld_nt(A.st.ctag + c);
        if (win_u)  // the window's rate of this cell: float(weight * scale), the product in double
            src_v[r] = A.src_weight ? (R)(A.src_weight[c] * scale_u) : R(0);
        else
            src_v[r] = su.src[c];
    }
    // ... and those of the west column and north row, RPW cells in the first lanes of every wave
    constexpr int RPW = 2 * TILE / (NTH / 64);
    const int lane = threadIdx.x & 63;
    const int rq = lane < RPW ? (int)(threadIdx.x >> 6) * RPW + lane : -1;
    const bool r_west = rq >= 0 && rq < TILE;
    const int r_j = rq < 0 ? 0 : (r_west ? rq : -1), r_i = rq < 0 ? 0 : (r_west ? -1 : rq - TILE);
    const int r_q = r_west ? 3 : 1;
    const unsigned r_c = rq < 0 ? 0u : cell_of(nb, r_j, r_i);
    const bool r_live = rq >= 0 && live[r_q];
    R r_res = R(0), r_src = R(0);
    float r_hc = 0.f;
    if (r_live) {
        r_res = su.res[r_c];
        r_hc = A.st.hc[r_c];
        r_src = win_u ? (A.src_weight ? (R)(A.src_weight[r_c] * scale_u) : R(0)) : su.src[r_c];
    }
    // 1. positivity ratio on cells [-2, 32]^2
    for (int p = threadIdx.x; p < 35 * 35; p += NTH) {
        const int jj = p / 35 - 2, ii = p - (p / 35) * 35 - 2;
        const int L = lds_of(jj, ii);
        const R out = pmax(s_qx[L + 1], R(0)) + pmax(-s_qx[L], R(0)) + pmax(s_qy[L + LS], R(0)) +
                      pmax(-s_qy[L], R(0));
        const R have = pmax(s_h[L], R(0)) * area;
        s_r[L] = out * dt_u > have ? dv<FAST>(have, out * dt_u) : R(1);
    }
    // 2. the own faces' momenta from the depths the last face step read, into LDS (the face loop reads them)
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const int L = lds_of(j, i);
        s_mx[j * TILE + i] = R(0.5) * (s_h[L - 1] + s_h[L]) * dx * s_u[L];
        s_my[j * TILE + i] = R(0.5) * (s_h[L - LS] + s_h[L]) * dx * s_v[L];
    }
    __syncthreads();
    // 3. limited discharges: x faces of cells [-1, 31] x [-1, 32], y faces of cells [-1, 32] x [-1, 31]
    for (int p = threadIdx.x; p < 33 * 34; p += NTH) {
        const int a = p / 34 - 1, b = p - (p / 34) * 34 - 1;
        const int Lx = lds_of(a, b);
        s_qx[Lx] = scaled(s_qx[Lx], s_r[Lx - 1], s_r[Lx]);
        const int Ly = lds_of(b, a);
        s_qy[Ly] = scaled(s_qy[Ly], s_r[Ly - LS], s_r[Ly]);
    }
    __syncthreads();
    // 4. the update of the own cells (three passes, as update_tile) ...
    bool wet = false;
    double src_sum = 0.0;
    {
        R hn_v[RO], resn_v[RO], s_v2[RO];
        int row_v[RO], mode_v[RO];
        float2 pair_v[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
            const int L = lds_of(j, i);
            const unsigned c = cs[r];
            const R qw = s_qx[L], qe = s_qx[L + 1], qn = s_qy[L], qs = s_qy[L + LS];
            const R h0 = s_h[L];
            const R inc = res_v[r] + dt_u * src_v[r] - dv<FAST>(dt_u * (qe - qw + qs - qn), area);
            const R sum = h0 + inc;
            const R hn = pmax(sum, R(0));
            resn_v[r] = sum > R(0) ? inc - (sum - h0) : R(0);
            hn_v[r] = hn;
            const R Hc = R(hc_v[r]);
            int mode = 0, row = 0;
            R sv = R(0);
            if (Hc > R(0)) {
                if (hn >= Hc) {
                    mode = 1;
                    row = A.st.last - 1;
                } else {
                    mode = 2;
                    sv = sq<FAST>(dv<FAST>(hn, Hc)) * R(A.st.last);
                    row = row_of(sv, A.st.last - 1);
                }
            }
            mode_v[r] = mode;
            row_v[r] = row;
            s_v2[r] = sv;
            if (mode) pair_v[r] = tag_v[r] == row ? A.st.cc[c] : A.st.ctab[(size_t)row * A.st.ncell + c];
        }
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
            const int L = lds_of(j, i);
            const unsigned c = cs[r];
            if (keep) {
                st_nt(s_qx[L], su.qx + c);
                st_nt(s_qy[L], su.qy + c);
            }
            if (win_u) {
                su.src[c] = src_v[r];
                src_sum += (double)src_v[r];
            }
            st_nt(resn_v[r], su.res_next + c);
            const R hn = hn_v[r];
            st_nt(hn, su.h_next + c);
            const R Hc = R(hc_v[r]);
            const float2 q = pair_v[r];
            R dn, plan = R(1);
            if (mode_v[r] == 0) {
                dn = hn;
            } else if (mode_v[r] == 1) {
                dn = R(q.y) + (hn - Hc);
            } else {
                const int row = row_v[r];
                if (plan_u) {
                    const R m = R(A.st.last);
                    plan = dv<FAST>(dv<FAST>(Hc * R(2 * row + 1), m * m), pmax(R(q.y) - R(q.x), R(1e-12)));
                }
                dn = R(q.x) + (s_v2[r] - R(row)) * (R(q.y) - R(q.x));
            }
            if (mode_v[r] && tag_v[r] != row_v[r]) {  // refill the cell's row cache
                A.st.cc[c] = q;
                A.st.ctag[c] = (unsigned char)row_v[r];
            }
            if (keep) st_nt(dn, su.d + c);
            if (plan_u) st_nt(plan, su.om + c);
            s_h[L] = hn;
            s_d[L] = dn;
            if (want_wave) s_r[L] = plan_u ? plan : su.om[c];
            wet |= hn > R(0);
        }
    }
    // ... and of the west column and north row
    if (rq >= 0) {
        const unsigned c = r_c;
        const int L = lds_of(r_j, r_i);
        if (r_live) {
            const R Hc = R(r_hc);
            const R qw = s_qx[L], qe = s_qx[L + 1], qn = s_qy[L], qs = s_qy[L + LS];
            const R h0 = s_h[L];
            const R inc = r_res + dt_u * r_src - dv<FAST>(dt_u * (qe - qw + qs - qn), area);
            const R sum = h0 + inc;
            const R hn = pmax(sum, R(0));
            const float2* tab = A.st.ctab + c;
            s_h[L] = hn;
            s_d[L] = cell_level<R, FAST>(hn, Hc, tab, A.st.ncell, A.st.last);
            if (want_wave) s_r[L] = plan_u ? cell_plan<R, FAST>(hn, Hc, tab, A.st.ncell, A.st.last) : su.om[c];
        } else {  // the stored state (its depth is in both buffers, so the staged one)
            s_d[L] = su.d[c];
            if (want_wave) s_r[L] = su.om[c];
        }
    }
    __syncthreads();
    // 5. boundary cells among the updated ones (their values after the step, as post_kernel sets them)
    if (flags & 1) {
        constexpr int NC = RO + 1;
        R hb[NC], db[NC];
        int Lb[NC];
        unsigned cb[NC];
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            Lb[r] = -1;
            int jj, ii, q;
            if (r < RO) {
                jj = (threadIdx.x >> 5) + r * (NTH / TILE);
                ii = i;
                q = 4;
            } else {
                if (rq < 0) continue;
                jj = r_j;
                ii = r_i;
                q = r_q;
            }
            if (!live[q]) continue;
            const unsigned c = cell_of(nb, jj, ii);
            const int e = G.bidx[c];
            if (e < 0) continue;
            const int kind = G.b_kind[e];
            if (kind == 0) {
                hb[r] = R(G.b_h0[e]);
                db[r] = R(G.b_d0[e]);
            } else if (kind == 1) {
                hb[r] = R(0);
                db[r] = R(0);
            } else {
                const int rf = G.b_ref[e], rt = rf / TILE2, rl = rf - rt * TILE2;
                int qr = -1;
                for (int k = 0; k < 9; ++k)
                    if (qr < 0 && nb[k] == rt && rt != 0) qr = k;
                if (qr < 0) continue;  // not among the updated cells: the run is set up so that this never happens
                const int Lr = lds_of((rl >> 5) + (qr / 3 - 1) * TILE, (rl & (TILE - 1)) + (qr % 3 - 1) * TILE);
                hb[r] = s_h[Lr];
                db[r] = pmax(s_d[Lr] + R(G.b_d0[e]), R(0));
            }
            Lb[r] = lds_of(jj, ii);
            cb[r] = c;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            if (Lb[r] < 0) continue;
            s_h[Lb[r]] = hb[r];
            s_d[Lb[r]] = db[r];
            if (r < RO) {
                su.h_next[cb[r]] = hb[r];
                if (keep) su.d[cb[r]] = db[r];
            }
        }
        __syncthreads();
    }
    // 6. running maxima, tally, wet flag and window source of the own cells
    if (sample) {
        const R time = (R)mb.u_time;
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
            const int L = lds_of(j, i);
            const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
            const R dk = s_d[L], hk = s_h[L];
            if (dk > A.mx_d[c]) A.mx_d[c] = dk;
            if (hk > A.mx_h[c]) A.mx_h[c] = hk;
            if (dk > A.deep) {
                const R uc = R(0.5) * (s_u[L] + s_u[L + 1]);
                const R vc = R(0.5) * (s_v[L] + s_v[L + LS]);
                const R sp = sqrt(uc * uc + vc * vc);
                if (sp > A.mx_s[c]) A.mx_s[c] = sp;
                if (sp * dk > A.mx_z[c]) A.mx_z[c] = sp * dk;
            }
            if (A.mx_t[c] < R(0) && dk >= A.arrive) A.mx_t[c] = time;
        }
    }
    if (flags & 2) {
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
            const int L = lds_of(j, i);
            const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
            const int2 tq = G.tidx[c];
            if (tq.x >= 0) {
                const double f = (double)(R(G.t_sign[tq.x]) * s_qx[L]);
                if (!failed) G.t_total[tq.x] += mb.u_dt * f;
                G.t_flow[tq.x] = f;
            }
            if (tq.y >= 0) {
                const double f = (double)(R(G.t_sign[tq.y]) * s_qy[L]);
                if (!failed) G.t_total[tq.y] += mb.u_dt * f;
                G.t_flow[tq.y] = f;
            }
        }
    }
    tile_flags<NTH>(t, wet, win_u, src_sum, &any_wet, red_src, A.wet, A.src_partial);
    if (!step) return;
    // 7. a face whose cell lies in a tile that is not listed keeps its stored discharge
    if (!(live[0] && live[1] && live[2] && live[3] && live[5] && live[6] && live[7] && live[8])) {
        for (int p = threadIdx.x; p < 33 * 34; p += NTH) {
            const int a = p / 34 - 1, b = p - (p / 34) * 34 - 1;
            if (!live[nb_of(a, b)]) s_qx[lds_of(a, b)] = su.qx[cell_of(nb, a, b)];
            if (!live[nb_of(b, a)]) s_qy[lds_of(b, a)] = su.qy[cell_of(nb, b, a)];
        }
        __syncthreads();
    }
    // 8. this step's faces
    const FaceIn<R> fx{s_u, s_v, s_qx, s_qy, s_h, s_d, 1, LS};
    const FaceIn<R> fy{s_v, s_u, s_qy, s_qx, s_h, s_d, LS, 1};
    const R dt_f = (R)mb.dt_step;
    R w = R(0);
#pragma unroll 1  // one cell at a time: 6 waves per SIMD without spilling
    for (int r = 0; r < RO; ++r) {
        const int j = (threadIdx.x >> 5) + r * (NTH / TILE);
        const int L = lds_of(j, i);
        const unsigned c = (unsigned)(t * TILE2 + j * TILE + i);
        const unsigned kw = west_of(t, j, i, nb), kn = north_of(t, j, i, nb);
        const R zxw = R(A.st.zcx[kw]), zxo = R(A.st.zcx[c]), zyn = R(A.st.zcy[kn]), zyo = R(A.st.zcy[c]);
        R omw = R(0), omn = R(0), omo = R(0);
        if (want_wave) {
            omw = s_r[L - 1];
            omn = s_r[L - LS];
            omo = s_r[L];
        }
        const R m = face_cell<R, FAST, false>(A, sf, fx, fy, L, c, s_mx[j * TILE + i], s_my[j * TILE + i], zxw, zxo,
                                              zyn, zyo, omw, omn, omo,
                                              dt_f, want_wave);
        w = m > w ? m : w;
    }
    if (want_wave) tile_wave<R, NTH>(A, t, member, w, red);
}

// Wet share of the listed tiles' cells (cells.cu:66-77).
template <class R, bool FAST>
__global__ __launch_bounds__(NT) void plan_kernel(const int* __restrict__ tiles, int n, const R* __restrict__ h,
                                                  R* __restrict__ om, const float* __restrict__ hc,
                                                  const float2* __restrict__ ctab, size_t ncell, int last) {
    if ((int)blockIdx.x >= n) return;
    const int t = tiles[blockIdx.x];
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const size_t c = (size_t)t * TILE2 + p;
        om[c] = cell_plan<R, FAST>(h[c], R(hc[c]), ctab + c, ncell, last);
    }
}

// Boundary cells (cells.cu:95-111): kind 0 fixed, 1 free (dry), 2 follows its inner cell lowered by d0.
template <class R>
__global__ void boundary_kernel(int n, const int* __restrict__ cell, const int* __restrict__ ref,
                                const int* __restrict__ kind, const float* __restrict__ d0,
                                const float* __restrict__ h0, R* __restrict__ h, R* __restrict__ d) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int k = cell[m];
    if (kind[m] == 0) {
        h[k] = R(h0[m]);
        d[k] = R(d0[m]);
    } else if (kind[m] == 1) {
        h[k] = R(0);
        d[k] = R(0);
    } else {
        h[k] = h[ref[m]];
        d[k] = pmax(d[ref[m]] + R(d0[m]), R(0));
    }
}

// Tiles that leave the active list (a protocol set_active) keep one mean depth and one residual in both buffers, at
// the handle's parity.
template <class R>
__global__ void leave_tiles_kernel(const int* __restrict__ tiles, int n, Bufs<R> B, int par) {
    if ((int)blockIdx.x >= n) return;
    const Cur<R> s = view(B, par);
    const size_t base = (size_t)tiles[blockIdx.x] * TILE2;
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        s.h_next[base + p] = s.h[base + p];
        s.res_next[base + p] = s.res[base + p];
    }
}

template <class R>
__device__ __forceinline__ void swap_at(R* a, R* b, size_t c) {
    const R x = a[c], y = b[c];
    a[c] = y;
    b[c] = x;
}

// Moves the buffered arrays of tiles [t0, t1) from parity `par` to `par2`: each buffer's content goes where the same
// role sits at the new parity (a swap, so a sleeping tile's second velocity buffer keeps its meaning), so that a member
// that took a different number of steps than the handle reads as the handle does.
template <class R>
__global__ void align_kernel(int t0, int t1, Bufs<R> B, int par, int par2) {
    if ((par & 1) == (par2 & 1)) return;
    const size_t c0 = (size_t)t0 * TILE2, c1 = (size_t)t1 * TILE2;
    for (size_t c = c0 + blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < c1; c += (size_t)gridDim.x * blockDim.x) {
        swap_at(B.h[0], B.h[1], c);
        swap_at(B.u[0], B.u[1], c);
        swap_at(B.v[0], B.v[1], c);
        swap_at(B.mx[0], B.mx[1], c);
        swap_at(B.my[0], B.my[1], c);
        swap_at(B.res[0], B.res[1], c);
    }
}

// The face momenta mx, my of the listed tiles (all active tiles: n, or glob->nact in a device-resident run) from the
// velocities and the mean depths the last step read, Abar * u as the face kernel recomputes them, into both buffers
// (a tile off the list keeps one momentum whichever buffer is current).
template <class R>
__global__ __launch_bounds__(NT) void momentum_kernel(const int* __restrict__ tiles, int n,
                                                      const RunGlobals* __restrict__ glob,
                                                      const int* __restrict__ nbr, const R* __restrict__ h_old,
                                                      const R* __restrict__ u, const R* __restrict__ v,
                                                      R* __restrict__ mx, R* __restrict__ my, R* __restrict__ mx2,
                                                      R* __restrict__ my2, R dx) {
    if ((int)blockIdx.x >= (glob ? glob->nact : n)) return;
    const int t = tiles[blockIdx.x];
    const int tw = nbr[(size_t)t * 8 + 3], tn = nbr[(size_t)t * 8 + 1];
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const int j = p / TILE, i = p - (p / TILE) * TILE;
        const size_t c = (size_t)t * TILE2 + p;
        const size_t w = i > 0 ? c - 1 : (size_t)tw * TILE2 + j * TILE + (TILE - 1);
        const size_t q = j > 0 ? c - TILE : (size_t)tn * TILE2 + (TILE - 1) * TILE + i;
        const R a = R(0.5) * (h_old[w] + h_old[c]) * dx * u[c];
        const R b = R(0.5) * (h_old[q] + h_old[c]) * dx * v[c];
        mx[c] = a, mx2[c] = a;
        my[c] = b, my2[c] = b;
    }
}

// Max of the active tiles' wave speeds into out[0], by one workgroup.
template <class R>
__global__ __launch_bounds__(1024) void wave_max_kernel(const int* __restrict__ active, int n,
                                                        const R* __restrict__ wave, double* __restrict__ out) {
    __shared__ double red[16];
    double m = 0.0;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        const double v = (double)wave[active[s]];
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = red[w] > r ? red[w] : r;
        out[0] = r;
    }
}

// One thread per face, each adding to its own FP64 total, no atomics (engine/maxima.py:53-62).
template <class R>
__global__ void tally_kernel(int n, double dt, const int* __restrict__ face, const int* __restrict__ axis,
                             const float* __restrict__ sign, const R* __restrict__ qx, const R* __restrict__ qy,
                             double* __restrict__ total, double* __restrict__ flow) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const R q = axis[m] ? qy[face[m]] : qx[face[m]];
    const double f = (double)(R(sign[m]) * q);
    total[m] += dt * f;
    flow[m] = f;
}

// Sum of doubles (or of h + res) by one workgroup into out[0]; deterministic order.
template <class R>
__global__ __launch_bounds__(1024) void sum_kernel(int64_t n, const int64_t* __restrict__ index,
                                                   const R* __restrict__ a, const R* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < n; p += blockDim.x) {
        if (x) {
            s += x[p];
        } else {
            const int64_t q = index ? index[p] : p;
            s += (double)a[q] + (b ? (double)b[q] : 0.0);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w];
        out[0] = r;
    }
}

// First stage of a deterministic sum: block b sums entries b, b + B, b + 2B, ... of (a[q] + b[q]) over q = index[p]
// (or p), in a fixed order; `partial` gets one double per block. The second stage is sum_kernel over `partial`.
constexpr int SUM_BLOCKS = 1024;
template <class R>
__global__ __launch_bounds__(256) void partial_sum_kernel(int64_t n, const int64_t* __restrict__ index,
                                                          const R* __restrict__ a, const R* __restrict__ b,
                                                          double* __restrict__ partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = index ? index[p] : p;
        s += (double)a[q] + (b ? (double)b[q] : 0.0);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Running maxima of the active tiles' own cells (engine/maxima.py:29-51).
template <class R>
__global__ __launch_bounds__(NT) void maxima_kernel(const int* __restrict__ active, int n, const int* __restrict__ nbr,
                                                    R time, R arrive, R deep, const R* __restrict__ d,
                                                    const R* __restrict__ h, const R* __restrict__ u,
                                                    const R* __restrict__ v, R* __restrict__ dmax, R* __restrict__ hmax,
                                                    R* __restrict__ smax, R* __restrict__ zmax, R* __restrict__ tarr) {
    if ((int)blockIdx.x >= n) return;
    const int t = active[blockIdx.x];
    const int te = nbr[(size_t)t * 8 + 4], ts = nbr[(size_t)t * 8 + 6];
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const int j = p / TILE, i = p - (p / TILE) * TILE;
        const size_t k = (size_t)t * TILE2 + p;
        const R dk = d[k];
        if (dk > dmax[k]) dmax[k] = dk;
        if (h[k] > hmax[k]) hmax[k] = h[k];
        if (dk > deep) {
            const size_t ke = i < TILE - 1 ? k + 1 : (size_t)te * TILE2 + j * TILE;
            const size_t ks = j < TILE - 1 ? k + TILE : (size_t)ts * TILE2 + i;
            const R uc = R(0.5) * (u[k] + u[ke]);
            const R vc = R(0.5) * (v[k] + v[ks]);
            const R s = sqrt(uc * uc + vc * vc);
            if (s > smax[k]) smax[k] = s;
            if (s * dk > zmax[k]) zmax[k] = s * dk;
        }
        if (tarr[k] < R(0) && dk >= arrive) tarr[k] = time;
    }
}

// Cells read by the host (w2d_read_cells): own cell at[q] into out[q].
template <class T>
__global__ void gather_kernel(int64_t n, const int64_t* __restrict__ at, const T* __restrict__ src, T* __restrict__ out) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = src[at[q]];
}

// The allocation-free gather (w2d_read_cells_noalloc): its index list and values live in `scratch`, a file-scope
// device array of NOALLOC_CELLS entries that each device holds for the life of the process. Layout: 256 bytes of
// flags (word 0: an index was not an own cell), then NOALLOC_CELLS int64 state indices, then NOALLOC_CELLS values.
constexpr int64_t NOALLOC_CELLS = 1 << 18;
constexpr size_t NOALLOC_FLAGS = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_FLAGS + (size_t)NOALLOC_CELLS * (sizeof(int64_t) + sizeof(double));
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

__global__ void noalloc_clear_kernel(int* flags) { flags[0] = 0; }

// State index k (the reference's halo layout) to the own-cell index, or -1 for a halo entry or one out of range;
// then the value of that own cell into out[q].
template <class T>
__global__ void noalloc_gather_kernel(int64_t n, int64_t nstate, const int64_t* __restrict__ k,
                                      const T* __restrict__ src, T* __restrict__ out, int* __restrict__ flags) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = k[q];
        int64_t own = -1;
        if (s >= 0 && s < nstate) {
            const int64_t t = s / STRIDE2, rem = s - t * STRIDE2;
            const int jj = (int)(rem / STRIDE), ii = (int)(rem - (rem / STRIDE) * STRIDE);
            if (jj >= HALO && jj < HALO + TILE && ii >= HALO && ii < HALO + TILE)
                own = t * TILE2 + (jj - HALO) * TILE + (ii - HALO);
        }
        if (own < 0) {
            atomicOr(flags, 1);
            out[q] = T(0);
        } else {
            out[q] = src[own];
        }
    }
}

// Cells written from the host (w2d_write_cells): value q into own cell at[q].
template <class T>
__global__ void scatter_kernel(int64_t n, const int64_t* __restrict__ at, const T* __restrict__ v, T* __restrict__ dst) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        dst[at[q]] = v[q];
}

// The window's source rate of every cell from its weight: float(weight * scale), the product in double.
template <class R>
__global__ void scale_source_kernel(size_t n, const double* __restrict__ w, double s, R* __restrict__ src) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
        src[p] = (R)(w[p] * s);
}

// ------------------------------------------------------------------------------------------------------------
// device-resident run: the time loop of wadi2d/solver/stepper.py:101-129 and the window loop of
// wadi2d/run.py:1712-1739, per member, with no host in the loop. A step is five launches: `pre_kernel` (one
// workgroup per member: the active-list check the last step asked for, the clock and source window, the list
// change they ask for, woken and leaving tiles), `gather_kernel` (the members' lists end to end, when one changed),
// the face and cell kernels, and `post_kernel` (one workgroup per member: its boundary cells, tally, wave-speed
// reduction, rain volume, clock and next step length). A member's tiles are one contiguous range and no tile has a
// neighbour in another member, so members never touch each other's tiles.

constexpr int SLEEP_CHECKS = 2;  // solver/stepper.py:14
constexpr int PRE_THREADS = 1024;

// Opens the member's next source window when its clock has reached the last one's end (run.py:1713-1733 with
// model.py:112-122: new rates, kept tiles, a forced wave refresh), decides this step's refresh and length
// (stepper.py:105-108), and retires a member that has reached its end or failed.
__device__ void ctl_begin(Member& b, const double* __restrict__ scales, const RunNum& num, double ex_sum, int& any) {
    b.stepping = 0;
    b.new_window = 0;
    b.refresh = 0;
    b.rebuild = 0;
    b.fuse = 0;
    if (b.done) return;
    if (b.t >= b.t_stop) return;  // held where the host asked (its end too): it reads this state before the next step
    // a failed member the host holds (a finite stop time: a job's chunks) stays listed as it failed, so the host
    // reads the failing step's state and samples its maxima over the same tiles as the host loop would; a failed
    // member whose last update waits for the pipe kernel is held for that step and retired after it
    if (b.failed && (b.t_stop < INFINITY || b.pend)) return;
    if (b.failed || b.t >= b.t_end - 1e-9) {
        b.done = 1;
        b.rebuild = 3;
        any = 1;
        return;
    }
    if (b.t >= b.window_end - 1e-9) {
        const int k = b.win;
        const double scale = k < b.nwin ? scales[b.scale_off + k] : 0.0;
        b.rain_done += b.rain_window;
        b.rain_window = 0.0;
        b.window_start = b.window_end;
        b.window_end = fmin(b.window_start + b.window_len, b.t_end);
        b.win = k + 1;
        if (num.excess) {
            // run.py:1719-1729 with forcing.py:667-678: the window's rain is the sum of its cells' excess rates
            // (excess_kernel), and a window opens when it rains or the last one did; the kept tiles are those with a
            // non-zero rate (the flags excess_kernel left)
            const double rain = ex_sum * num.area * (b.window_end - b.window_start);
            if (rain > 0.0 || !b.dry) {
                b.scale = 1.0;
                b.new_window = 1;
                b.refresh_due = 1;
                b.rebuild = 1;
                any = 1;
            }
            b.rain_window = rain;
            b.dry = !(rain > 0.0);
        } else {
            if (scale != 0.0 || !b.dry) {
                b.scale = scale;
                b.new_window = 1;
                b.refresh_due = 1;
                b.rebuild = 1;
                any = 1;
            }
            b.dry = !(scale > 0.0);
        }
    }
    b.refresh = b.refresh_due || (b.nsteps % num.dt_every) == 0;
    b.refresh_due = 0;
    b.dt_step = fmin(fmin(b.dt, b.window_end - b.t), num.dt_max);
    b.stepping = 1;
    // this step's update is left to the next step's pipe kernel unless the next step reads the wet flags it leaves (a
    // check follows a step whose count is a multiple of active_every) or may change the list or the source (a window
    // end, the run's end or a hold: the clock after this step is post_kernel's t + dt_step, the same sum). Unless the
    // next step cannot change anything this update leaves. A member with every tile listed: a window's end can only
    // add tiles, so it changes nothing (but a step that opened its window books that window's rain after its update);
    // a check changes nothing when every tile is also kept whatever its wet flag says (a source tile while it rains,
    // a boundary tile). The excess source rewrites the cells' rates, and the tiles it keeps, before the next step.
    if (b.pipe) {
        const double t1 = b.t + b.dt_step;
        const int own = b.tile1 - b.tile0;
        const bool listed = !num.excess && b.nlist == own;
        const bool window_kept = listed && !b.new_window;
        const bool check_kept = listed && (b.scale != 0.0 ? b.kept_rain : b.kept_dry) == own;
        b.fuse = (check_kept || (b.nsteps % num.active_every) != 0) && (window_kept || !(t1 >= b.window_end - 1e-9)) &&
                 !(t1 >= b.t_end - 1e-9) && !(t1 >= b.t_stop);
    }
}

// Whether ctl_begin opens a window of this member before its next step (the clock has reached the window's end and
// the member goes on).
__device__ __forceinline__ bool window_due(const Member& b) {
    return !b.done && !b.failed && b.t < b.t_end - 1e-9 && b.t < b.t_stop && b.t >= b.window_end - 1e-9;
}

// The SCS cumulative excess of depth p, mm (hydrology/losses.py:59-72): S = (1000 / cn - 10) 25.4, Ia = 0.2 S,
// max(p - Ia, 0)^2 / (max(p - Ia, 0) + S), read through numpy's nan_to_num (forcing.py:659, 664).
__device__ __forceinline__ double scs_excess(double p, double cn) {
    const double s = (1000.0 / cn - 10.0) * 25.4;
    const double ia = 0.2 * s;
    const double x = p - ia;
    const double eff = (x >= 0.0 || isnan(x)) ? x : 0.0;  // np.maximum(x, 0.0)
    const double v = eff * eff / (eff + s);
    if (isnan(v)) return 0.0;
    if (isinf(v)) return v > 0.0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
    return v;
}

// One cell's rate over a window whose storm depth runs from d0 to d1 mm (forcing.py:644-672): its depth is the
// storm's times the cell's factor w (a depth grid; 1 without one), its excess that depth where it has no curve
// number (cn NaN), and the rate float((E1 - E0) / denom), denom = 1000 (t1 - t0).
template <class R>
__device__ __forceinline__ R excess_rate(double w, double cn, double d0, double d1, double denom) {
    double e0 = d0 * w, e1 = d1 * w;
    if (!isnan(cn)) {
        e0 = scs_excess(e0, cn);
        e1 = scs_excess(e1, cn);
    }
    return (R)((e1 - e0) / denom);
}

constexpr int EX_THREADS = 256;

// The excess source of whole tiles: every cell's rate into src, and per tile the FP64 sum of its rates and whether
// one is non-zero. In a device-resident run (mem set) only the tiles of members that open a window before their
// next step, with that window's (d0, d1, denom) from terms; in a protocol call (mem null) every tile, with the
// given window. Each tile is one workgroup's; a tile's cells are all rated, listed or not (forcing.py:667-672 sets
// every interior cell's rate). On the steps where no member's window is due (most), every workgroup leaves after
// reading the M members' clocks. With gridded rain (pix set) a cell's storm depths are its pixel's, pix[c] indexing
// a row of npix depths at the window's start followed by npix at its end (pdep: the protocol call's row, or the
// run's rows, one per window in the order of the terms); a cell with no pixel (-1) takes none.
template <class R>
__global__ __launch_bounds__(EX_THREADS) void excess_kernel(int nt, const Member* __restrict__ mem, int nmem,
                                                          const int* __restrict__ tile_member,
                                                          const double* __restrict__ terms, double d0a, double d1a,
                                                          double dna, const double* __restrict__ w,
                                                          const double* __restrict__ cn, const int* __restrict__ pix,
                                                          const double* __restrict__ pdep, int npix,
                                                          R* __restrict__ src, double* __restrict__ partial,
                                                          int* __restrict__ flag) {
    __shared__ double red[EX_THREADS / 64];
    __shared__ int any_nz, any_due;
    if (mem) {
        if (threadIdx.x == 0) any_due = 0;
        __syncthreads();
        for (int m = threadIdx.x; m < nmem; m += EX_THREADS)
            if (window_due(mem[m])) any_due = 1;
        __syncthreads();
        if (!any_due) return;
    }
    for (int t = 1 + blockIdx.x; t < nt; t += gridDim.x) {
        double d0 = d0a, d1 = d1a, dn = dna;
        const double* prow = pdep;
        if (mem) {
            const Member& b = mem[tile_member[t]];
            if (t < b.tile0 || t >= b.tile1 || !window_due(b)) continue;
            if (b.win < b.nwin) {
                const double* q = terms + 3 * (b.scale_off + b.win);
                d0 = q[0], d1 = q[1], dn = q[2];
                if (pix) prow = pdep + (size_t)2 * npix * (size_t)(b.scale_off + b.win);
            } else {
                d0 = 0.0, d1 = 0.0, dn = 1.0;
                prow = nullptr;
            }
        }
        if (threadIdx.x == 0) any_nz = 0;
        __syncthreads();
        double sum = 0.0;
        int nz = 0;
#pragma unroll
        for (int r = 0; r < TILE2 / EX_THREADS; ++r) {
            const size_t c = (size_t)t * TILE2 + r * EX_THREADS + threadIdx.x;
            double c0 = d0, c1 = d1;
            if (pix) {
                const int q = pix[c];
                c0 = q >= 0 && prow ? prow[q] : 0.0;
                c1 = q >= 0 && prow ? prow[npix + q] : 0.0;
            }
            const R v = excess_rate<R>(w[c], cn[c], c0, c1, dn);
            src[c] = v;
            sum += (double)v;
            nz |= v != R(0);
        }
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
        if (nz) any_nz = 1;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = red[0];
            for (int q = 1; q < EX_THREADS / 64; ++q) tot += red[q];
            partial[t] = tot;
            flag[t] = any_nz;
        }
        __syncthreads();
    }
}

// Block-wide exclusive prefixes of three flags per thread over a chunk of PRE_THREADS, with one barrier pair:
// returns this thread's slots and adds the chunk's counts to `base` (in every thread).
__device__ __forceinline__ int3 chunk_slots(bool f0, bool f1, bool f2, int (*wave_cnt)[PRE_THREADS / 64], int3& base) {
    const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1), m2 = __ballot(f2);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    if (lane == 0) {
        wave_cnt[0][wv] = __popcll(m0);
        wave_cnt[1][wv] = __popcll(m1);
        wave_cnt[2][wv] = __popcll(m2);
    }
    __syncthreads();
    int3 off = make_int3(0, 0, 0), total = make_int3(0, 0, 0);
    for (int q = 0; q < PRE_THREADS / 64; ++q) {
        const int c0 = wave_cnt[0][q], c1 = wave_cnt[1][q], c2 = wave_cnt[2][q];
        if (q < wv) {
            off.x += c0;
            off.y += c1;
            off.z += c2;
        }
        total.x += c0;
        total.y += c1;
        total.z += c2;
    }
    __syncthreads();
    const int3 slot = make_int3(base.x + off.x + __popcll(m0 & below), base.y + off.y + __popcll(m1 & below),
                                base.z + off.z + __popcll(m2 & below));
    base.x += total.x;
    base.y += total.y;
    base.z += total.z;
    return slot;
}

struct RunLists {
    const int* nbr;
    const int* wet;
    const int* has_source;
    const int* bnd_tile;
    int* is_active;
    int* idle;
    int* flags;
    int* seg;      // per member, its active tiles from seg[tile0]
    int* woken;    // per member, from woken[tile0]
    int* leaving;  // per member, from leaving[tile0]
    int* active;
};

// One member's active tiles (stepper.py:56-92) by `mode`: 1 keeps what is listed and adds wet and kept tiles with
// their neighbours (retain); 2 does that and lets a tile go after SLEEP_CHECKS checks that did not want it (check);
// 3 drops the member's tiles (done). The list stays in ascending order; woken and leaving tiles go to their lists.
// One workgroup, the member's own tiles.
constexpr int SHARED_TILES = 32768;  // a member with no more tiles keeps its tile flags in LDS, one byte each

// SH: the flags in LDS (`sf`, indexed from tile0): bit 0 wanted, bit 1 active after, bit 2 active before; else in
// L.flags (bits 0 and 1), the before flag read again from is_active. Loops over the member's tiles are unrolled so
// that several tiles' loads are in flight at once.
template <bool SH>
__device__ void rebuild_flags(Member& b, int mode, const RunLists& L, int (*wave_cnt)[PRE_THREADS / 64], int& n_woken,
                              int& n_leaving, unsigned char* __restrict__ sf) {
    const int t0 = b.tile0, t1 = b.tile1;
    const bool look = mode == 1 || mode == 2;
    const bool raining = b.scale != 0.0;
#pragma unroll 4
    for (int t = t0 + threadIdx.x; t < t1; t += PRE_THREADS) {
        int wanted = 0;
        if (look) {
            const int keep = (L.has_source[t] && raining) || L.bnd_tile[t];
            wanted = L.wet[t] || keep;
        }
        if constexpr (SH)
            sf[t - t0] = (unsigned char)wanted;
        else
            L.flags[t] = wanted;
    }
    __syncthreads();
#pragma unroll 2
    for (int t = t0 + threadIdx.x; t < t1; t += PRE_THREADS) {
        const int cur = L.is_active[t];
        int a = cur;
        if (mode == 3) {
            a = 0;
        } else if (look) {
            // neighbours are this member's tiles or the void tile, whose flag is never set
            int nb[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) nb[q] = L.nbr[(size_t)t * 8 + q];
            if constexpr (SH) {
                a = sf[t - t0] & 1;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    a |= (unsigned)(nb[q] - t0) < (unsigned)(t1 - t0) ? sf[nb[q] - t0] & 1 : 0;
            } else {
                a = L.flags[t] & 1;
#pragma unroll
                for (int q = 0; q < 8; ++q) a |= L.flags[nb[q]] & 1;
            }
            if (mode == 1) {
                a |= cur;
            } else {
                L.idle[t] = a ? 0 : L.idle[t] + 1;
                a |= cur && L.idle[t] < SLEEP_CHECKS;
            }
        }
        if constexpr (SH)
            sf[t - t0] |= (unsigned char)((a << 1) | (cur << 2));
        else
            L.flags[t] |= a << 1;
    }
    __syncthreads();
    int3 n = make_int3(0, 0, 0);  // active, woken, leaving
    for (int base = t0; base < t1; base += PRE_THREADS) {
        const int t = base + threadIdx.x;
        const bool in = t < t1;
        int a, cur;
        if constexpr (SH) {
            const int f = in ? sf[t - t0] : 0;
            a = (f >> 1) & 1;
            cur = (f >> 2) & 1;
        } else {
            a = in ? (L.flags[t] >> 1) & 1 : 0;
            cur = in ? L.is_active[t] : 0;
        }
        const bool wk = a && !cur && look;
        const bool lv = cur && !a;
        const int3 slot = chunk_slots(a, wk, lv, wave_cnt, n);
        if (a) L.seg[t0 + slot.x] = t;
        if (wk) L.woken[t0 + slot.y] = t;
        if (lv) L.leaving[t0 + slot.z] = t;
        if (in) L.is_active[t] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) b.nlist = n.x;
    n_woken = n.y;
    n_leaving = n.z;
    __syncthreads();
}

__device__ void rebuild(Member& b, int mode, const RunLists& L, int (*wave_cnt)[PRE_THREADS / 64], int& n_woken,
                        int& n_leaving, unsigned char* sf) {
    if (b.tile1 - b.tile0 <= SHARED_TILES)
        rebuild_flags<true>(b, mode, L, wave_cnt, n_woken, n_leaving, sf);
    else
        rebuild_flags<false>(b, mode, L, wave_cnt, n_woken, n_leaving, sf);
}

// What a change of the list does to tiles (stepper.py:94-99): a woken tile's wet share from its depth, by the
// kernel's own arithmetic; a tile that leaves keeps one mean depth in both buffers, and a finished member's tile also
// one velocity.
template <class R, bool FAST>
__device__ void transitions(const Member& b, int nw, int nl, const RunLists& L, const Bufs<R>& B,
                            const float* __restrict__ hc, const float2* __restrict__ ctab, size_t ncell, int last,
                            R dx, int fresh) {
    const Cur<R> s = view(B, b.par);
    // tiles are taken TB at a time, every load of a batch issued before its stores (a member that finishes moves all
    // its tiles here at once)
    constexpr int TB = 4;
    const int p = threadIdx.x, j = p / TILE, i = p - (p / TILE) * TILE;
    for (int q0 = 0; q0 < nw; q0 += TB) {
        R hv[TB], hcv[TB];
        size_t cv[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            cv[u] = q0 + u < nw ? (size_t)L.woken[b.tile0 + q0 + u] * TILE2 + p : 0;
            hv[u] = q0 + u < nw ? s.h[cv[u]] : R(0);
            hcv[u] = q0 + u < nw ? R(hc[cv[u]]) : R(0);
        }
#pragma unroll
        for (int u = 0; u < TB; ++u)
            if (q0 + u < nw) s.om[cv[u]] = cell_plan<R, FAST>(hv[u], hcv[u], ctab + cv[u], ncell, last);
    }
    // a leaving tile's face momentum is stored (both buffers) from the depths the last step read, before those
    // depth buffers are made one; a stored momentum (set from the host since) is already in both
    if (nl && !fresh) {
        const int p2 = b.par & 1;
        for (int q0 = 0; q0 < nl; q0 += TB) {
            R av[TB], ev[TB];
            size_t cv[TB];
#pragma unroll
            for (int u = 0; u < TB; ++u) {
                if (q0 + u >= nl) continue;
                const int t = L.leaving[b.tile0 + q0 + u];
                const size_t c = (size_t)t * TILE2 + p;
                const size_t w = i > 0 ? c - 1 : (size_t)L.nbr[(size_t)t * 8 + 3] * TILE2 + j * TILE + (TILE - 1);
                const size_t n = j > 0 ? c - TILE : (size_t)L.nbr[(size_t)t * 8 + 1] * TILE2 + (TILE - 1) * TILE + i;
                cv[u] = c;
                av[u] = R(0.5) * (s.h_next[w] + s.h_next[c]) * dx * s.u[c];
                ev[u] = R(0.5) * (s.h_next[n] + s.h_next[c]) * dx * s.v[c];
            }
#pragma unroll
            for (int u = 0; u < TB; ++u) {
                if (q0 + u >= nl) continue;
                B.mx[p2][cv[u]] = av[u], B.mx[1 - p2][cv[u]] = av[u];
                B.my[p2][cv[u]] = ev[u], B.my[1 - p2][cv[u]] = ev[u];
            }
        }
        __syncthreads();
    }
    // then one mean depth and one residual in both buffers, and for a finished member one velocity (the momentum is in
    // both already)
    for (int q0 = 0; q0 < nl; q0 += TB) {
        R hv[TB], rv[TB], uv[TB], vv[TB];
        size_t cv[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            if (q0 + u >= nl) continue;
            cv[u] = (size_t)L.leaving[b.tile0 + q0 + u] * TILE2 + p;
            hv[u] = s.h[cv[u]];
            rv[u] = s.res[cv[u]];
            if (b.done) uv[u] = s.u[cv[u]], vv[u] = s.v[cv[u]];
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            if (q0 + u >= nl) continue;
            s.h_next[cv[u]] = hv[u];
            s.res_next[cv[u]] = rv[u];
            if (b.done) s.u1[cv[u]] = uv[u], s.v1[cv[u]] = vv[u];
        }
    }
    __syncthreads();
}

// The members' lists end to end, in member (so ascending tile) order, when one of them changed; one workgroup.
// Members are taken PRE_THREADS at a time: a block scan of their list lengths gives each its offset, then every thread
// copies entries of the chunk's joined lists, finding an entry's member by bisection over the offsets, several
// entries' loads in flight at once (one member of 16,641 tiles and 64 of 25,136 in all alike).
__device__ void gather_lists(const Member* __restrict__ mem, int M, RunGlobals* __restrict__ g,
                             const int* __restrict__ seg, int* __restrict__ active,
                             int (*wave_cnt)[PRE_THREADS / 64]) {
    __shared__ int s_off[PRE_THREADS + 1], s_t0[PRE_THREADS];
    if (!g->changed) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int base = 0;  // entries of the members before this chunk
    for (int m0 = 0; m0 < M; m0 += PRE_THREADS) {
        const int m = m0 + (int)threadIdx.x;
        const int n = m < M ? mem[m].nlist : 0;
        const int t0 = m < M ? mem[m].tile0 : 0;
        int x = n;  // inclusive scan within the wave
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wave_cnt[0][wv] = x;
        __syncthreads();
        int before = 0, total = 0;
        for (int q = 0; q < PRE_THREADS / 64; ++q) {
            const int c = wave_cnt[0][q];
            if (q < wv) before += c;
            total += c;
        }
        s_off[threadIdx.x] = before + x - n;
        s_t0[threadIdx.x] = t0;
        if (threadIdx.x == 0) s_off[PRE_THREADS] = total;
        __syncthreads();
        const int cnt = M - m0 < PRE_THREADS ? M - m0 : PRE_THREADS;
        constexpr int U = 4;
        for (int e0 = (int)threadIdx.x; e0 < total; e0 += U * PRE_THREADS) {
            int val[U], at[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * PRE_THREADS;
                at[u] = -1;
                if (e >= total) continue;
                int lo = 0, hi = cnt - 1;  // the last member whose offset is <= e
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                val[u] = seg[s_t0[lo] + (e - s_off[lo])];
                at[u] = base + e;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (at[u] >= 0) active[at[u]] = val[u];
        }
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        g->nact = base;
        g->changed = 0;
    }
}

// Before a step, one workgroup per member: the check the last step asked for (stepper.py:113-114), the clock and
// window, the list change those ask for, and the tiles that woke or left; the last workgroup to finish then gathers
// the members' lists into the active list.
template <class R, bool FAST>
__global__ __launch_bounds__(PRE_THREADS) void pre_kernel(Member* __restrict__ mem, RunGlobals* __restrict__ g,
                                                          const double* __restrict__ scales, const RunNum num,
                                                          RunLists L, Bufs<R> B, const float* __restrict__ hc,
                                                          const float2* __restrict__ ctab, size_t ncell, int last,
                                                          R dx, int fresh, unsigned long long* __restrict__ member_wave,
                                                          const double* __restrict__ src_partial, int flush_only) {
    static_assert(PRE_THREADS == TILE2, "transitions give one thread to each cell of a tile");
    __shared__ int wave_cnt[3][PRE_THREADS / 64];
    __shared__ int any;
    __shared__ unsigned char sflags[SHARED_TILES];
    Member& b = mem[blockIdx.x];
    int nw, nl;
    bool changed = false;
    // with every tile of the member listed, a window's rebuild (which keeps what is listed) changes nothing, and so
    // does a check when every tile is also kept whatever its wet flag says: it only sets each tile's idle count to 0
    const int own = b.tile1 - b.tile0;
    const bool listed = b.nlist == own;
    if (b.check) {
        if (listed && !num.excess && (b.scale != 0.0 ? b.kept_rain : b.kept_dry) == own) {
            for (int t = b.tile0 + threadIdx.x; t < b.tile1; t += PRE_THREADS) L.idle[t] = 0;
        } else {
            rebuild(b, 2, L, wave_cnt, nw, nl, sflags);
            transitions<R, FAST>(b, nw, nl, L, B, hc, ctab, ncell, last, dx, fresh);
            changed = true;
        }
        if (threadIdx.x == 0) b.check = 0;
    }
    if (!flush_only) {
        // the window's rain in the excess form: the sum of the member's tile sums excess_kernel left
        __shared__ double ex_red[PRE_THREADS / 64];
        __shared__ int ex_due;
        if (threadIdx.x == 0) ex_due = num.excess && window_due(b);
        __syncthreads();
        double ex_sum = 0.0;
        if (ex_due) {
            for (int t = b.tile0 + threadIdx.x; t < b.tile1; t += PRE_THREADS) ex_sum += src_partial[t];
            for (int off = 32; off > 0; off >>= 1) ex_sum += __shfl_xor(ex_sum, off);
            if ((threadIdx.x & 63) == 0) ex_red[threadIdx.x >> 6] = ex_sum;
            __syncthreads();
            ex_sum = ex_red[0];
            for (int q = 1; q < PRE_THREADS / 64; ++q) ex_sum += ex_red[q];
        }
        if (threadIdx.x == 0) {
            int mine = 0;
            ctl_begin(b, scales, num, ex_sum, mine);
            if (b.refresh) member_wave[blockIdx.x] = 0ull;
            if (b.stepping && !b.fuse) g->classic = 1;
            any = mine;
        }
        __syncthreads();
        if (any && !(b.rebuild == 1 && b.nlist == own)) {
            rebuild(b, b.rebuild, L, wave_cnt, nw, nl, sflags);
            transitions<R, FAST>(b, nw, nl, L, B, hc, ctab, ncell, last, dx, fresh);
            changed = true;
        }
    }
    if (changed && threadIdx.x == 0) g->changed = 1;
    // the last workgroup to finish gathers the members' lists (every other one has published its own first); a
    // lone workgroup is the last without the handshake (its device-scope fences and atomic)
    __shared__ int is_last;
    const bool alone = gridDim.x == 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (alone) {
            is_last = 1;
        } else {
            __threadfence();
            is_last = atomicAdd(&g->pre_done, 1) == (int)gridDim.x - 1;
        }
    }
    __syncthreads();
    if (!is_last) return;
    if (!alone) __threadfence();
    gather_lists(mem, (int)gridDim.x, g, L.seg, L.active, wave_cnt);
    if (threadIdx.x == 0 && !alone) g->pre_done = 0;
}

// After a step, one workgroup per member: its boundary cells (cells.cu:95-111) and FP64 tally
// (engine/maxima.py:53-62), the wave-speed reduction and the next step length with the failure rules
// (stepper.py:109-126), the window's rain volume, the clock, and the check every active_every steps.
constexpr int POST_THREADS = 1024;  // a member's tiles, boundary cells and tally faces are strided over them

template <class R>
__global__ __launch_bounds__(POST_THREADS) void post_kernel(Member* __restrict__ mem,
                                                   const unsigned long long* __restrict__ member_wave,
                                                   const int* __restrict__ is_active,
                                                   const double* __restrict__ src_partial, const RunNum num,
                                                   const int* __restrict__ cell, const int* __restrict__ ref,
                                                   const int* __restrict__ kind, const float* __restrict__ d0,
                                                   const float* __restrict__ h0, Bufs<R> B,
                                                   const int* __restrict__ face, const int* __restrict__ axis,
                                                   const float* __restrict__ sign, double* __restrict__ total,
                                                   double* __restrict__ flow, int drain, RunGlobals* g) {
    __shared__ double red_s[POST_THREADS / 64];
    __shared__ double red_q[POST_THREADS / 64];
    if (blockIdx.x == 0 && threadIdx.x == 0) g->classic = 0;
    Member& b = mem[blockIdx.x];
    const bool pend = b.pend != 0;
    const bool step = !drain && b.stepping;
    if (!pend && !step) return;
    // the window's rain of a step that opened one: the source sums of the member's listed tiles
    auto rain_sum = [&]() {
        double rs = 0.0;
        for (int t = b.tile0 + threadIdx.x; t < b.tile1; t += blockDim.x)
            if (is_active[t]) rs += src_partial[t];
        for (int off = 32; off > 0; off >>= 1) rs += __shfl_xor(rs, off);
        if ((threadIdx.x & 63) == 0) red_s[threadIdx.x >> 6] = rs;
    };
    auto reduced = [&](const double* red) {
        double v = red[0];
        for (int q = 1; q < POST_THREADS / 64; ++q) v += red[q];
        return v;
    };
    if (pend) {
        // the last step's update, just made by the pipe kernel with its boundary cells and tally: its peak outflow
        // from the flows it left, and its window's rain
        double out_q = 0.0;
        for (int q = b.f0 + threadIdx.x; q < b.f1; q += blockDim.x) out_q += flow[q];
        for (int off = 32; off > 0; off >>= 1) out_q += __shfl_xor(out_q, off);
        if ((threadIdx.x & 63) == 0) red_q[threadIdx.x >> 6] = out_q;
        const bool books_rain = b.u_new_window && !num.excess;
        if (books_rain) rain_sum();
        __syncthreads();
        if (threadIdx.x == 0) {
            if (books_rain) b.rain_window = reduced(red_s) * num.area * (b.window_end - b.window_start);
            if (b.f1 > b.f0 && !b.failed) {
                const double q = reduced(red_q);
                if (q > b.q_peak) {
                    b.q_peak = q;
                    b.t_peak = b.u_time;
                }
            }
            b.pend = 0;
            if (!drain) b.u_made += 1;  // the drain's update follows the last step
        }
        __syncthreads();
        if (!step) return;
    }
    const bool fused = b.fuse != 0;  // this step's update waits for the next step's pipe kernel
    // the step's outputs: the next buffers of the parities it stepped at (thread 0 advances them below)
    const Cur<R> s = view(B, b.par);
    R* const h = s.h_next;
    R* const d = s.d;
    if (!fused)
        for (int p = b.b0 + threadIdx.x; p < b.b1; p += blockDim.x) {
            const int k = cell[p];
            if (kind[p] == 0) {
                h[k] = R(h0[p]);
                d[k] = R(d0[p]);
            } else if (kind[p] == 1) {
                h[k] = R(0);
                d[k] = R(0);
            } else {
                h[k] = h[ref[p]];
                d[k] = pmax(d[ref[p]] + R(d0[p]), R(0));
            }
        }
    // the step's wave speed and whether it fails (stepper.py:117-124), known to every thread before the books: a
    // failing step books nothing, as the host loop raises before its after-step work (tally, smallest dt, peak)
    const double w = b.refresh ? wave_value<R>(member_wave[blockIdx.x]) : 0.0;
    const double dt_next = b.refresh ? fmin(num.cfl * num.dx / fmax(w, num.wave_floor), num.dt_max) : b.dt;
    const int failing = !b.refresh ? 0 : !isfinite(w) ? 1 : dt_next < num.dt_min ? 2 : 0;
    double out_q = 0.0;  // this thread's share of the member's discharge out, summed in a fixed order
    if (!fused)
        for (int q = b.f0 + threadIdx.x; q < b.f1; q += blockDim.x) {
            const R qq = axis[q] ? s.qy[face[q]] : s.qx[face[q]];
            const double f = (double)(R(sign[q]) * qq);
            if (!failing) total[q] += b.dt_step * f;
            flow[q] = f;
            out_q += f;
        }
    for (int off = 32; off > 0; off >>= 1) out_q += __shfl_xor(out_q, off);
    if ((threadIdx.x & 63) == 0) red_q[threadIdx.x >> 6] = out_q;
    // the excess form's rain is booked by the pre kernel; a waiting update's by the next post kernel
    const bool books_rain = b.new_window && !num.excess && !fused;
    if (books_rain) rain_sum();
    __syncthreads();  // every thread has read the parities before thread 0 moves them on
    if (threadIdx.x != 0) return;
    if (books_rain) b.rain_window = reduced(red_s) * num.area * (b.window_end - b.window_start);
    const long long before = b.nsteps;
    if (fused) {  // what the pipe kernel needs of this step when it makes its update
        b.pend = 1;
        b.u_dt = b.dt_step;
        b.u_want = b.refresh;
        b.u_new_window = b.new_window;
        b.u_scale = b.scale;
        b.u_par = b.par;
        b.u_nsteps = before;
        b.u_time = b.t + b.dt_step;
    }
    b.t += b.dt_step;
    b.nsteps = before + 1;
    if (b.f1 > b.f0 && !failing && !fused) {  // the peak outflow over the member's steps (the run summary's, monitor.py)
        const double q = reduced(red_q);
        if (q > b.q_peak) {
            b.q_peak = q;
            b.t_peak = b.t;
        }
    }
    if (failing != 1) b.dt = dt_next;  // a non-finite speed leaves dt as it was (stepper.py:117-119)
    b.failed = failing;
    if (!failing) b.dt_min = fmin(b.dt_min, b.dt);
    if (before % num.active_every == 0) b.check = 1;
    b.par = 1 - b.par;
}

// Running maxima after the step, at the member's new time (engine/maxima.py:29-51), on the listed tiles: those with
// a boundary cell, whose values the post kernel sets after the cell kernel (the cell kernel samples every other
// tile).
template <class R>
__global__ __launch_bounds__(NT) void maxima_run_kernel(const int* __restrict__ tiles, const int* __restrict__ is_active,
                                                        const Member* __restrict__ mem,
                                                        const int* __restrict__ tile_member, const int* __restrict__ nbr,
                                                        R arrive, R deep, Bufs<R> B, R* __restrict__ dmax,
                                                        R* __restrict__ hmax, R* __restrict__ smax,
                                                        R* __restrict__ zmax, R* __restrict__ tarr, int every) {
    const int t = uni(tiles[blockIdx.x]);
    if (!is_active[t]) return;
    const Member& b = mem[uni(tile_member[t])];
    // sampled after the steps whose count is a multiple of `every`, as the host loop's on_step does; the member's
    // parities have moved on, so its current buffers hold the step's results
    if (!b.stepping || b.fuse || b.nsteps % every != 0) return;  // a waiting update samples in the pipe kernel
    const Cur<R> s = view(B, uni(b.par));
    const R* __restrict__ d = s.d;
    const R* __restrict__ h = s.h;
    const R* __restrict__ u = s.u;
    const R* __restrict__ v = s.v;
    const R time = (R)b.t;
    const int te = nbr[(size_t)t * 8 + 4], ts = nbr[(size_t)t * 8 + 6];
    for (int p = threadIdx.x; p < TILE2; p += NT) {
        const int j = p / TILE, i = p - (p / TILE) * TILE;
        const size_t k = (size_t)t * TILE2 + p;
        const R dk = d[k];
        if (dk > dmax[k]) dmax[k] = dk;
        if (h[k] > hmax[k]) hmax[k] = h[k];
        if (dk > deep) {
            const size_t ke = i < TILE - 1 ? k + 1 : (size_t)te * TILE2 + j * TILE;
            const size_t ks = j < TILE - 1 ? k + TILE : (size_t)ts * TILE2 + i;
            const R uc = R(0.5) * (u[k] + u[ke]);
            const R vc = R(0.5) * (v[k] + v[ks]);
            const R sp = sqrt(uc * uc + vc * vc);
            if (sp > smax[k]) smax[k] = sp;
            if (sp * dk > zmax[k]) zmax[k] = sp * dk;
        }
        if (tarr[k] < R(0) && dk >= arrive) tarr[k] = time;
    }
}

// A plain copy, the yardstick for the memory counters and for the bandwidth a stream reaches on the part: n float4
// read once and written once, one per thread (a grid-stride loop reaches about 4.8 TB/s on MI355X, this 6.2).
__global__ __launch_bounds__(256) void copy_probe_kernel(size_t n, const float4* __restrict__ a, float4* __restrict__ b) {
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p < n) b[p] = a[p];
}

template <class R>
__global__ void fill_kernel(R* a, size_t n, R v) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) a[p] = v;
}

#include "levels_kernels.hpp"

}  // namespace

// ------------------------------------------------------------------------------------------------------------
// host side

struct w2d_handle {
    // the learned closure (off while closure_n is 0)
    int closure_n = 0, closure_conveyance = 0;
    int *cl_cell_r = nullptr, *cl_cell_l = nullptr, *cl_axis = nullptr;
    float *cl_stat = nullptr, *cl_norm = nullptr, *cl_params = nullptr;
    uint16_t* cl_wbf = nullptr;
    float2 *corr_x = nullptr, *corr_y = nullptr;
    int prec = 4;
    int fast = 0;
    int face_cache = 1;
    int faces_threads = 512;
    int cells_threads = 256;
    int ntiles = 0;
    int rows = 0;
    size_t ncell = 0;
    Params P{};
    hipStream_t stream = nullptr;
    // device state, element size prec: the buffers of Bufs, and the handle's parity (see Cur); in a device-resident
    // run each member keeps its own, brought back to this one by run_align
    void *hb[2] = {}, *ub[2] = {}, *vb[2] = {}, *mxb[2] = {}, *myb[2] = {};
    void *d = nullptr, *qx = nullptr, *qy = nullptr, *om = nullptr;
    int par = 0;
    // the face momentum: steps do not store it (the face kernel recomputes it from u and the depths the last step
    // read); m_fresh says whether mx, my hold it, as after set_field or materialize; it goes to the kernels by value
    int m_fresh = 1;
    void *qx1b[2] = {}, *qy1b[2] = {}, *resb[2] = {}, *src = nullptr;
    void* wave = nullptr;
    int* wet = nullptr;
    // statics
    float* hc = nullptr;
    float2* ctab = nullptr;
    float4 *ftabx = nullptr, *ftaby = nullptr, *sx = nullptr, *sy = nullptr;
    float2 *es = nullptr, *cwv = nullptr;
    float *zcx = nullptr, *zcy = nullptr;
    float4 *fcx = nullptr, *fcy = nullptr;
    float2* cc = nullptr;
    unsigned char *ftagx = nullptr, *ftagy = nullptr, *ctag = nullptr;
    int* nbr = nullptr;
    std::vector<int32_t> nbr_host;
    // lists
    int* active = nullptr;
    int nact = 0;
    std::vector<int32_t> active_host;
    int* tiles_scratch = nullptr;
    int nb_cells = 0;
    int *b_cell = nullptr, *b_ref = nullptr, *b_kind = nullptr;
    std::vector<int32_t> b_cell_host, b_ref_host, b_kind_host, t_face_host, t_axis_host;
    // the pipelined step of device-resident runs (w2d_run_pipeline): asked for, taken by the run set up last, and its
    // per-tile flags and per-cell boundary and tally indices
    int pipe_on = 1;
    int run_pipe = 0;
    std::vector<int32_t> pipe_ok;  // per member of the run set up last: its boundary cells allow the pipelined step
    int* pipe_near = nullptr;
    int* pipe_bidx = nullptr;
    int2* pipe_tidx = nullptr;
    float *b_d0 = nullptr, *b_h0 = nullptr;
    int n_tally = 0;
    int *t_face = nullptr, *t_axis = nullptr;
    float* t_sign = nullptr;
    double *t_total = nullptr, *t_flow = nullptr;
    int64_t n_store = 0;
    int64_t* store_index = nullptr;
    void* maxima[5] = {};
    double m_arrive = 0.1, m_deep = 0.05;
    double* scalar = nullptr;  // reduction results
    double* src_weight = nullptr;
    double* partial = nullptr;  // SUM_BLOCKS doubles
    // device-resident run
    int run_m = 0;
    Member* run_mem = nullptr;
    RunGlobals* run_glob = nullptr;
    RunNum run_num{};
    int *run_tile_member = nullptr, *run_has_source = nullptr, *run_bnd_tile = nullptr, *run_is_active = nullptr,
        *run_idle = nullptr, *run_flags = nullptr, *run_woken = nullptr, *run_leaving = nullptr, *run_seg = nullptr;
    int* run_bnd_list = nullptr;  // the tiles with a boundary cell (run_nbnd of them), for maxima_run_kernel
    int run_nbnd = 0;
    double *run_scales = nullptr, *run_src_partial = nullptr;
    unsigned long long* run_wave = nullptr;  // per member, the step's largest wave speed as bits
    int run_maxima = 0;
    // local time stepping (w2d_level_*): ratio, old depth and level, edge estimates, per-tile levels; the lists of
    // the current plan; the smallest step the halos are read at
    void *lv_r = nullptr, *lv_h0 = nullptr, *lv_d0 = nullptr, *lv_est = nullptr;
    int* lv_level = nullptr;
    std::vector<int*> lv_lists;
    std::vector<int> lv_count, lv_width;
    int lv_sub = 0;
    void* cells_buf = nullptr;  // w2d_read_cells: the indices and values of a gather, grown as needed
    size_t cells_cap = 0;
    // steps were enqueued since the members' parities were last brought back to the handle's (run_align), or the
    // handle stepped on its own since the members took its parities
    int run_dirty = 0, members_stale = 0;
    std::vector<double> weight_host;
    // the excess form of the source (w2d_source_losses): src_weight holds depth factors, src_cn curve numbers
    int excess = 0;
    double* src_cn = nullptr;
    std::vector<double> run_terms_host;  // (d0, d1, denom) per window of a device-resident run
    double* run_terms = nullptr;
    double* ex_partial = nullptr;  // per tile, a protocol call's sums and flags
    int* ex_flag = nullptr;
    // gridded rain: each own cell's pixel (-1: none), a protocol call's pixel depths, a run's per-window rows
    int* src_pix = nullptr;
    int npix = 0;
    double* pix_dep = nullptr;
    size_t pix_dep_n = 0;
    std::vector<double> run_pix_host;
    double* run_pix_terms = nullptr;
    // the device this handle was created on, and its copy of the allocation-free gather's scratch array (the
    // address hipGetSymbolAddress gave on that device: the array's name is not a device address in host code)
    int device = 0;
    void* scratch_dev = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;

    // kernel timing: when on, every step records events around its two kernels on the handle's stream
    int timing = 0;
    std::vector<hipEvent_t> event_pool;
    std::vector<hipEvent_t> marks;  // triples: before faces, after faces, after cells
    std::vector<void*> owned;
};

namespace {

size_t esize(const w2d_handle* H) { return H->prec == 8 ? sizeof(double) : sizeof(float); }

int dalloc(w2d_handle* H, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? 2 : 3, std::string("hipMalloc: ") + hipGetErrorString(e));
    H->owned.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes ? bytes : 16, H->stream);
    if (e != hipSuccess) return fail(3, std::string("hipMemset: ") + hipGetErrorString(e));
    return 0;
}

// Each device's address of `scratch`, looked up once per device, with that device current, at the first handle's
// set-up on it (w2d_create); every later use takes this address, never the array's name.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[MAX_DEVICES] = {};

int scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(1, "scratch: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(3, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

template <class T>
int upload(w2d_handle* H, T** dst, const T* src, size_t count) {
    int s = dalloc(H, (void**)dst, count * sizeof(T));
    if (s) return s;
    HIP_TRY(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

template <class R>
Bufs<R> bufs(const w2d_handle* H) {
    Bufs<R> b;
    for (int k = 0; k < 2; ++k) {
        b.h[k] = (R*)H->hb[k];
        b.u[k] = (R*)H->ub[k];
        b.v[k] = (R*)H->vb[k];
        b.mx[k] = (R*)H->mxb[k];
        b.my[k] = (R*)H->myb[k];
        b.res[k] = (R*)H->resb[k];
        b.qx1[k] = (R*)H->qx1b[k];
        b.qy1[k] = (R*)H->qy1b[k];
    }
    b.d = (R*)H->d;
    b.qx = (R*)H->qx;
    b.qy = (R*)H->qy;
    b.om = (R*)H->om;
    b.src = (R*)H->src;
    return b;
}

// The handle's current view of its buffers (the element type does not matter for the pointers).
Cur<char> cur(const w2d_handle* H) { return view(bufs<char>(H), H->par); }

// Device pointer of field f (the current buffer for the buffered ones), or null for a field the device keeps no copy of.
void* field_ptr(w2d_handle* H, int f) {
    const Cur<char> c = cur(H);
    switch (f) {
        case W2D_H: return c.h;
        case W2D_D: return c.d;
        case W2D_U: return c.u;
        case W2D_V: return c.v;
        case W2D_MX: return c.mx;
        case W2D_MY: return c.my;
        case W2D_QX: return c.qx;
        case W2D_QY: return c.qy;
        case W2D_U1: return c.u1;
        case W2D_V1: return c.v1;
        case W2D_MX1: return H->mxb[1 - (H->par & 1)];
        case W2D_MY1: return H->myb[1 - (H->par & 1)];
        case W2D_QX1: return c.qx1;
        case W2D_QY1: return c.qy1;
        case W2D_SRC: return H->src;
        case W2D_RES: return c.res;
        case W2D_OM: return c.om;
        default: return nullptr;
    }
}

// Every buffer of field f a set_field writes, so that each holds the one value the reference keeps (the momentum and
// the mean depth); the velocities' two buffers are the reference's own.
std::vector<void*> field_all(w2d_handle* H, int f) {
    switch (f) {
        case W2D_H: return {H->hb[0], H->hb[1]};
        case W2D_RES: return {H->resb[0], H->resb[1]};
        case W2D_QX1: return {H->qx1b[0], H->qx1b[1]};
        case W2D_QY1: return {H->qy1b[0], H->qy1b[1]};
        case W2D_MX:
        case W2D_MX1: return {H->mxb[0], H->mxb[1]};
        case W2D_MY:
        case W2D_MY1: return {H->myb[0], H->myb[1]};
        default: {
            void* p = field_ptr(H, f);
            if (!p) return {};
            return {p};
        }
    }
}

template <class R>
StepArgs<R> step_args(w2d_handle* H, double dt, int want_wave) {
    StepArgs<R> A{};
    A.b = bufs<R>(H);
    A.par = H->par;
    A.st.hc = H->hc;
    A.st.ctab = H->ctab;
    A.st.ftabx = H->ftabx;
    A.st.ftaby = H->ftaby;
    A.st.sx = H->sx;
    A.st.sy = H->sy;
    A.st.es = H->es;
    A.st.cwv = H->cwv;
    A.st.zcx = H->zcx;
    A.st.zcy = H->zcy;
    A.st.fcx = H->face_cache ? H->fcx : nullptr;
    A.st.fcy = H->face_cache ? H->fcy : nullptr;
    A.corr_x = H->closure_n ? H->corr_x : nullptr;
    A.corr_y = H->closure_n ? H->corr_y : nullptr;
    A.st.ftagx = H->ftagx;
    A.st.ftagy = H->ftagy;
    A.st.cc = H->cc;
    A.st.ctag = H->ctag;
    A.st.nbr = H->nbr;
    A.st.ncell = H->ncell;
    A.st.last = H->rows - 1;
    A.P = H->P;
    A.active = H->active;
    A.nact = H->nact;
    // the reference hands float32 kernels np.float32(dt) (gpu/backend.py:153)
    A.dt = (R)dt;
    A.want_wave = want_wave;
    A.wave = (R*)H->wave;
    A.wet = H->wet;
    A.fresh = H->m_fresh;
    return A;
}

int mark(w2d_handle* H) {
    hipEvent_t e;
    if (H->event_pool.empty()) {
        HIP_TRY(hipEventCreate(&e));
    } else {
        e = H->event_pool.back();
        H->event_pool.pop_back();
    }
    HIP_TRY(hipEventRecord(e, H->stream));
    H->marks.push_back(e);
    return 0;
}

// One step's kernels on `blocks` workgroups (a device-resident run launches one per tile; those without a listed tile
// leave at once): the learned closure when it is on, then the face kernel and the cell kernel at the handle's
// workgroup sizes (256 or 512 threads). With timing on, marks go before the first kernel, between the two and after.
template <class R, bool FAST>
int launch_step_kernels(w2d_handle* H, const StepArgs<R>& A, int blocks) {
    int s;
    if (H->timing && (s = mark(H))) return s;
    const bool learned = H->closure_n > 0;
    if (learned) {
        if constexpr (std::is_same<R, float>::value) {
            ClosureArgs C{H->closure_n, H->cl_cell_r, H->cl_cell_l, H->cl_axis, H->cl_stat, H->cl_norm, H->cl_wbf,
                          H->cl_params, H->sx, H->sy, H->corr_x, H->corr_y, H->closure_conveyance, A.members,
                          A.tile_member};
            if (A.members || A.want_wave) {
                const int cb = std::min((H->closure_n + 63) / 64, 4096);
                closure_kernel<<<cb, 256, 0, H->stream>>>(C, A.b, A.par);
                HIP_TRY(hipGetLastError());
            }
            faces_kernel<float, FAST, 512, true><<<blocks, 512, 0, H->stream>>>(A);
        } else {
            return fail(1, "the learned closure runs in FP32");
        }
    } else if (H->faces_threads == 512) {
        faces_kernel<R, FAST, 512><<<blocks, 512, 0, H->stream>>>(A);
    } else {
        faces_kernel<R, FAST, 256><<<blocks, 256, 0, H->stream>>>(A);
    }
    HIP_TRY(hipGetLastError());
    if (H->timing && (s = mark(H))) return s;
    if (H->cells_threads == 512)
        cells_kernel<R, FAST, 512><<<blocks, 512, 0, H->stream>>>(A);
    else
        cells_kernel<R, FAST, 256><<<blocks, 256, 0, H->stream>>>(A);
    HIP_TRY(hipGetLastError());
    if (H->timing && (s = mark(H))) return s;
    return 0;
}

template <class R, bool FAST>
int launch_step(w2d_handle* H, double dt, int want_wave) {
    const StepArgs<R> A = step_args<R>(H, dt, want_wave);
    int s;
    if ((s = launch_step_kernels<R, FAST>(H, A, H->nact))) return s;
    if (H->nb_cells) {
        const Cur<R> c = view(A.b, H->par);
        boundary_kernel<R><<<(H->nb_cells + 255) / 256, 256, 0, H->stream>>>(
            H->nb_cells, H->b_cell, H->b_ref, H->b_kind, H->b_d0, H->b_h0, c.h_next, c.d);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// Stores the face momentum of every listed tile (both buffers) if steps have run since it was last stored, so that
// whatever then changes u, v, h or the list, mx and my keep the value the last step left.
int materialize(w2d_handle* H) {
    if (H->m_fresh) return 0;
    const bool run = H->run_m > 0;
    const int blocks = run ? H->ntiles : H->nact;
    if (blocks > 0) {
        const Cur<char> c = cur(H);
        const int p2 = H->par & 1;
        if (H->prec == 8)
            momentum_kernel<double><<<blocks, NT, 0, H->stream>>>(
                H->active, H->nact, run ? H->run_glob : nullptr, H->nbr, (const double*)c.h_next, (const double*)c.u,
                (const double*)c.v, (double*)H->mxb[p2], (double*)H->myb[p2], (double*)H->mxb[1 - p2],
                (double*)H->myb[1 - p2], H->P.dx);
        else
            momentum_kernel<float><<<blocks, NT, 0, H->stream>>>(
                H->active, H->nact, run ? H->run_glob : nullptr, H->nbr, (const float*)c.h_next, (const float*)c.u,
                (const float*)c.v, (float*)H->mxb[p2], (float*)H->myb[p2], (float*)H->mxb[1 - p2],
                (float*)H->myb[1 - p2], (float)H->P.dx);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(H->stream));
    H->m_fresh = 1;
    return 0;
}

// After steps of a device-resident run: applies the list check the last step asked for, brings every member's buffers
// to one parity (the first member's) and takes it as the handle's, and follows the device's active list. Every call that
// reads or writes the state through the protocol makes it first (a no-op when no run step was enqueued since).
int run_align(w2d_handle* H);

// Fields that hold the face momentum, and fields whose change would move the momentum the last step left.
bool momentum_store(int f) { return f == W2D_MX || f == W2D_MY || f == W2D_MX1 || f == W2D_MY1; }
bool momentum_field(int f) { return momentum_store(f) || f == W2D_H || f == W2D_U || f == W2D_V; }

int run_step(w2d_handle* H, double dt, int want_wave) {
    if (H->prec == 8) return launch_step<double, false>(H, dt, want_wave);
    if (H->fast) return launch_step<float, true>(H, dt, want_wave);
    return launch_step<float, false>(H, dt, want_wave);
}

// own-cell index (t * 1024 + j * 32 + i) of a state index, or -1 for a halo entry or one out of range
int64_t own_of(const w2d_handle* H, int64_t k) {
    if (k < 0 || k >= (int64_t)H->ntiles * STRIDE2) return -1;
    const int64_t t = k / STRIDE2;
    const int64_t rem = k - t * STRIDE2;
    const int jj = (int)(rem / STRIDE), ii = (int)(rem - (rem / STRIDE) * STRIDE);
    if (jj < HALO || jj >= HALO + TILE || ii < HALO || ii >= HALO + TILE) return -1;
    return t * TILE2 + (jj - HALO) * TILE + (ii - HALO);
}

// state layout (with halos) from the own-cell array: every halo entry from its owning tile, the void's zeros
template <class T>
void expand(const w2d_handle* H, const T* own, T* out) {
    const int nt = H->ntiles;
    for (int t = 0; t < nt; ++t) {
        for (int jj = 0; jj < STRIDE; ++jj)
            for (int ii = 0; ii < STRIDE; ++ii) {
                const int j = jj - HALO, i = ii - HALO;
                const int dj = j < 0 ? -1 : (j >= TILE ? 1 : 0);
                const int di = i < 0 ? -1 : (i >= TILE ? 1 : 0);
                int tt = t;
                if (dj || di) {
                    const int q = (dj + 1) * 3 + di + 1;
                    tt = t == 0 ? 0 : H->nbr_host[(size_t)t * 8 + (q < 4 ? q : q - 1)];
                }
                out[(size_t)t * STRIDE2 + jj * STRIDE + ii] =
                    own[(size_t)tt * TILE2 + (j - dj * TILE) * TILE + (i - di * TILE)];
            }
    }
}

template <class T>
void compact(const w2d_handle* H, const T* state, T* own) {
    for (int t = 0; t < H->ntiles; ++t)
        for (int j = 0; j < TILE; ++j)
            for (int i = 0; i < TILE; ++i)
                own[(size_t)t * TILE2 + j * TILE + i] = state[(size_t)t * STRIDE2 + (j + HALO) * STRIDE + i + HALO];
}

// sum of a[q] (+ b[q]) over q in index (or 0 .. n-1), FP64, deterministic: many blocks, then one.
template <class R>
int big_sum(w2d_handle* H, int64_t n, const int64_t* index, const R* a, const R* b, double* out) {
    partial_sum_kernel<R><<<SUM_BLOCKS, 256, 0, H->stream>>>(n, index, a, b, H->partial);
    HIP_TRY(hipGetLastError());
    sum_kernel<float><<<1, 1024, 0, H->stream>>>(SUM_BLOCKS, nullptr, nullptr, nullptr, H->partial, H->scalar);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, H->scalar, sizeof(double), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

}  // namespace

// Shared with the library's other translation units (csrc/mlp.hip, csrc/tables_gpu.hip): record the message w2d_last_error returns, and
// return `code`.
int w2d_report(int code, const char* msg) { return fail(code, msg); }

extern "C" {

const char* w2d_last_error(void) { return g_error.c_str(); }

int w2d_device_count(int32_t* out) {
    if (!out) return fail(1, "device_count: bad arguments");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    return 0;
}

int w2d_create(const w2d_tables* T, const w2d_params* p, int32_t precision, int32_t variant, w2d_handle** out) {
    if (!T || !p || !out || T->ntiles < 1 || T->rows < 2 || (precision != W2D_F32 && precision != W2D_F64))
        return fail(1, "w2d_create: bad arguments");
    auto* H = new w2d_handle();
    *out = nullptr;
    H->prec = precision;
    H->fast = precision == W2D_F32 && (variant & 1) == W2D_FAST;
    // the pipelined step by default in FP32; in FP64 its LDS (80 KB a workgroup) leaves one workgroup per CU and the
    // two-kernel step is faster (it can still be asked for)
    H->pipe_on = precision == W2D_F32;
    H->face_cache = (variant & W2D_NO_FACE_CACHE) == 0;
    H->ntiles = T->ntiles;
    H->rows = T->rows;
    H->ncell = (size_t)T->ntiles * TILE2;
    if (H->ncell * 16 > (size_t)INT32_MAX) {
        delete H;
        return fail(1, "w2d_create: more cells than 32-bit cell offsets reach (at most 134,217,728)");
    }
    H->P = Params{p->g, p->dx, p->e_wet, p->h_cfl, p->u_max, p->lead, p->a_min, T->rows, p->second, p->sheet,
                  p->mirror, p->carry};
    int s = 0;
    auto bail = [&](int code) {
        w2d_destroy(H);
        return code;
    };
    if (hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(3, "hipStreamCreate failed"));
    if (hipGetDevice(&H->device) != hipSuccess) return bail(fail(3, "hipGetDevice failed"));
    if ((s = scratch_address(H->device, &H->scratch_dev))) return bail(s);
    const size_t n = H->ncell, es = esize(H);
    for (void** pair : {H->hb, H->ub, H->vb, H->mxb, H->myb, H->resb, H->qx1b, H->qy1b})
        for (int b = 0; b < 2; ++b)
            if ((s = dalloc(H, &pair[b], n * es))) return bail(s);
    for (void** a : {&H->d, &H->qx, &H->qy, &H->om, &H->src})
        if ((s = dalloc(H, a, n * es))) return bail(s);
    if ((s = dalloc(H, &H->wave, (size_t)H->ntiles * es))) return bail(s);
    if ((s = dalloc(H, (void**)&H->wet, (size_t)H->ntiles * sizeof(int)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->active, (size_t)H->ntiles * sizeof(int)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->tiles_scratch, (size_t)H->ntiles * sizeof(int)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->scalar, 16 * sizeof(double)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->partial, SUM_BLOCKS * sizeof(double)))) return bail(s);

    // statics into the device layout
    const int rows = T->rows, last = rows - 1;
    {
        std::vector<float2> ctab((size_t)last * n);
        for (size_t c = 0; c < n; ++c)
            for (int r = 0; r < last; ++r) ctab[(size_t)r * n + c] = make_float2(T->dtab[c * rows + r], T->dtab[c * rows + r + 1]);
        if ((s = upload(H, &H->ctab, ctab.data(), ctab.size()))) return bail(s);
    }
    for (int axis = 0; axis < 2; ++axis) {
        const float* ft = axis ? T->ftaby : T->ftabx;
        std::vector<float4> f4((size_t)last * n);
        for (size_t c = 0; c < n; ++c)
            for (int r = 0; r < last; ++r) {
                const float* q = ft + (c * rows + r) * 2;
                f4[(size_t)r * n + c] = make_float4(q[0], q[1], q[2], q[3]);
            }
        if ((s = upload(H, axis ? &H->ftaby : &H->ftabx, f4.data(), f4.size()))) return bail(s);
    }
    {
        std::vector<float4> sx(n), sy(n);
        std::vector<float2> es(n), cwv(n);
        // a face's table top carries in its sign whether the face has a crest coefficient (exact: a sign flip);
        // the coefficient itself is read only for such a face
        auto flag = [](float ef, float cw) { return cw != 0.0f ? -std::fabs(ef) : std::fabs(ef); };
        for (size_t c = 0; c < n; ++c) {
            sx[c] = make_float4(T->flx[c], T->frx[c], T->glx[c], T->grx[c]);
            sy[c] = make_float4(T->fly[c], T->fry[c], T->gly[c], T->gry[c]);
            es[c] = make_float2(flag(T->efx[c], T->cwx[c]), flag(T->efy[c], T->cwy[c]));
            cwv[c] = make_float2(T->cwx[c], T->cwy[c]);
        }
        for (size_t c = 0; c < n; ++c)
            if (!(T->efx[c] >= 0.0f) || !(T->efy[c] >= 0.0f)) return bail(fail(1, "create: a face table top is negative or NaN"));
        if ((s = upload(H, &H->sx, sx.data(), n))) return bail(s);
        if ((s = upload(H, &H->sy, sy.data(), n))) return bail(s);
        if ((s = upload(H, &H->es, es.data(), n))) return bail(s);
        if ((s = upload(H, &H->cwv, cwv.data(), n))) return bail(s);
    }
    if ((s = upload(H, &H->hc, T->hc, n))) return bail(s);
    H->nbr_host.assign(T->nbr, T->nbr + (size_t)H->ntiles * 8);
    for (int q = 0; q < 8; ++q) H->nbr_host[q] = 0;  // the void tile's neighbours are the void
    if ((s = upload(H, &H->nbr, H->nbr_host.data(), H->nbr_host.size()))) return bail(s);
    // centre-line beds in the own layout; a face reads its upwind neighbour's from that tile (the void's is 0),
    // which is what static_cell_state's filled halos hold (state.py:216-223)
    if ((s = upload(H, &H->zcx, T->zcx, n))) return bail(s);
    if ((s = upload(H, &H->zcy, T->zcy, n))) return bail(s);
    // row caches, empty
    if ((s = dalloc(H, (void**)&H->fcx, n * sizeof(float4)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->fcy, n * sizeof(float4)))) return bail(s);
    if ((s = dalloc(H, (void**)&H->cc, n * sizeof(float2)))) return bail(s);
    for (unsigned char** tg : {&H->ftagx, &H->ftagy, &H->ctag}) {
        if ((s = dalloc(H, (void**)tg, n))) return bail(s);
        if (hipMemsetAsync(*tg, 0xFF, n, H->stream) != hipSuccess) return bail(fail(3, "create: tag fill"));
    }
    std::vector<int32_t> all;
    for (int t = 1; t < H->ntiles; ++t) all.push_back(t);
    if ((s = w2d_set_active(H, all.data(), (int)all.size()))) return bail(s);
    if (hipStreamSynchronize(H->stream) != hipSuccess) return bail(fail(3, "create: device error"));
    *out = H;
    return 0;
}

int w2d_destroy(w2d_handle* H) {
    if (!H) return 0;
    if (H->stream) (void)hipStreamSynchronize(H->stream);
    for (hipEvent_t e : H->marks) (void)hipEventDestroy(e);
    for (hipEvent_t e : H->event_pool) (void)hipEventDestroy(e);
    if (H->graph) (void)hipGraphExecDestroy(H->graph);
    for (void* p : H->owned) (void)hipFree(p);
    if (H->stream) (void)hipStreamDestroy(H->stream);
    delete H;
    return 0;
}

int w2d_set_field(w2d_handle* H, int32_t f, const void* host, int64_t n) {
    if (!H || !host || n != (int64_t)H->ntiles * STRIDE2) return fail(1, "set_field: bad arguments");
    void* dst = field_ptr(H, f);
    if (!dst) return fail(1, "set_field: the device keeps no copy of this field");
    int s;
    if ((s = run_align(H))) return s;
    if (momentum_field(f) && (s = materialize(H))) return s;
    const size_t es = esize(H);
    std::vector<char> own(H->ncell * es);
    if (es == 8)
        compact(H, (const double*)host, (double*)own.data());
    else
        compact(H, (const float*)host, (float*)own.data());
    // the void tile stays zero whatever the host wrote into tile 0
    std::memset(own.data(), 0, TILE2 * es);
    // every buffer of a field the reference keeps once (the momentum: one value per face whichever buffer is current)
    for (void* b : field_all(H, f)) HIP_TRY(hipMemcpyAsync(b, own.data(), own.size(), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_get_field(w2d_handle* H, int32_t f, void* host, int64_t n) {
    if (!H || !host || n != (int64_t)H->ntiles * STRIDE2) return fail(1, "get_field: bad arguments");
    void* src = field_ptr(H, f);
    if (!src) return fail(1, "get_field: the device keeps no copy of this field (the positivity ratio lives in LDS)");
    int s;
    if ((s = run_align(H))) return s;
    src = field_ptr(H, f);
    if (momentum_store(f) && (s = materialize(H))) return s;
    const size_t es = esize(H);
    std::vector<char> own(H->ncell * es);
    HIP_TRY(hipMemcpyAsync(own.data(), src, own.size(), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    if (es == 8)
        expand(H, (const double*)own.data(), (double*)host);
    else
        expand(H, (const float*)own.data(), (float*)host);
    return 0;
}

int w2d_read_cells(w2d_handle* H, int32_t f, const int64_t* index, int64_t n, void* out) {
    if (!H || n < 0 || (n && (!index || !out))) return fail(1, "read_cells: bad arguments");
    void* src = field_ptr(H, f);
    if (!src) return fail(1, "read_cells: the device keeps no copy of this field");
    int s;
    if ((s = run_align(H))) return s;
    src = field_ptr(H, f);
    if (momentum_store(f) && (s = materialize(H))) return s;
    if (n == 0) return 0;
    const size_t es = esize(H);
    std::vector<int64_t> at(n);
    for (int64_t m = 0; m < n; ++m) {
        at[m] = own_of(H, index[m]);
        if (at[m] < 0) return fail(1, "read_cells: an index is not an own cell");
    }
    // one gather on the device and one copy back, whatever the number of cells
    const size_t need = n * (sizeof(int64_t) + es);
    if (H->cells_cap < need) {
        if (H->cells_buf) {
            H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), H->cells_buf), H->owned.end());
            (void)hipFree(H->cells_buf);
            H->cells_buf = nullptr;
            H->cells_cap = 0;
        }
        if ((s = dalloc(H, &H->cells_buf, need))) return s;
        H->cells_cap = need;
    }
    int64_t* d_at = (int64_t*)H->cells_buf;
    void* d_out = (char*)H->cells_buf + n * sizeof(int64_t);
    HIP_TRY(hipMemcpyAsync(d_at, at.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, H->stream));
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    if (es == 8)
        gather_kernel<double><<<blocks, 256, 0, H->stream>>>(n, d_at, (const double*)src, (double*)d_out);
    else
        gather_kernel<float><<<blocks, 256, 0, H->stream>>>(n, d_at, (const float*)src, (float*)d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d_out, n * es, hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_read_cells_noalloc(w2d_handle* H, int32_t f, const int64_t* index, int64_t n, void* out) {
    if (!H || n < 0 || (n && (!index || !out))) return fail(1, "read_cells_noalloc: bad arguments");
    if (n > NOALLOC_CELLS)
        return fail(1, "read_cells_noalloc: more than 262,144 cells in one call (the scratch array's size)");
    void* src = field_ptr(H, f);
    if (!src) return fail(1, "read_cells_noalloc: the device keeps no copy of this field");
    if (!H->scratch_dev) return fail(3, "read_cells_noalloc: no scratch address for this handle's device");
    int s;
    if ((s = run_align(H))) return s;
    src = field_ptr(H, f);
    if (momentum_store(f) && (s = materialize(H))) return s;
    if (n == 0) return 0;
    const size_t es = esize(H);
    unsigned char* base = (unsigned char*)H->scratch_dev;
    int* d_flags = (int*)base;
    int64_t* d_at = (int64_t*)(base + NOALLOC_FLAGS);
    void* d_val = base + NOALLOC_FLAGS + (size_t)NOALLOC_CELLS * sizeof(int64_t);
    noalloc_clear_kernel<<<1, 1, 0, H->stream>>>(d_flags);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d_at, index, n * sizeof(int64_t), hipMemcpyHostToDevice, H->stream));
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    const int64_t nstate = (int64_t)H->ntiles * STRIDE2;
    if (es == 8)
        noalloc_gather_kernel<double><<<blocks, 256, 0, H->stream>>>(n, nstate, d_at, (const double*)src,
                                                                      (double*)d_val, d_flags);
    else
        noalloc_gather_kernel<float><<<blocks, 256, 0, H->stream>>>(n, nstate, d_at, (const float*)src,
                                                                     (float*)d_val, d_flags);
    HIP_TRY(hipGetLastError());
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(out, d_val, n * es, hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipMemcpyAsync(&bad, d_flags, sizeof(int), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    if (bad) return fail(1, "read_cells_noalloc: an index is not an own cell");
    return 0;
}

int w2d_write_cells(w2d_handle* H, int32_t f, const int64_t* index, int64_t n, const void* values) {
    if (!H || n < 0 || (n && (!index || !values))) return fail(1, "write_cells: bad arguments");
    void* dst = field_ptr(H, f);
    if (!dst) return fail(1, "write_cells: the device keeps no copy of this field");
    if (n == 0) return 0;
    const size_t es = esize(H);
    // own cells of real tiles; an index given twice keeps its last value, as an assignment through it would
    std::vector<std::pair<int64_t, int64_t>> cm(n);  // (own cell, position in the call)
    for (int64_t m = 0; m < n; ++m) {
        const int64_t c = own_of(H, index[m]);
        if (c < (int64_t)TILE2) return fail(1, "write_cells: an index is not an own cell of a tile");
        cm[m] = {c, m};
    }
    std::sort(cm.begin(), cm.end());
    std::vector<int64_t> at;
    std::vector<char> val;
    for (size_t q = 0; q < cm.size(); ++q) {
        if (q + 1 < cm.size() && cm[q + 1].first == cm[q].first) continue;  // a later write of the same cell
        const int64_t m = cm[q].second;
        at.push_back(cm[q].first);
        val.insert(val.end(), (const char*)values + m * es, (const char*)values + (m + 1) * es);
    }
    int s;
    if ((s = run_align(H))) return s;
    if (momentum_field(f) && (s = materialize(H))) return s;
    const int64_t k = (int64_t)at.size();
    void* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, k * (sizeof(int64_t) + es)));
    int64_t* d_at = (int64_t*)buf;
    void* d_val = (char*)buf + k * sizeof(int64_t);
    const std::vector<void*> targets = field_all(H, f);
    hipError_t e = hipMemcpyAsync(d_at, at.data(), k * sizeof(int64_t), hipMemcpyHostToDevice, H->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_val, val.data(), k * es, hipMemcpyHostToDevice, H->stream);
    const unsigned blocks = (unsigned)std::min<int64_t>((k + 255) / 256, 1024);
    for (void* t : targets) {
        if (e != hipSuccess) break;
        if (es == 8)
            scatter_kernel<double><<<blocks, 256, 0, H->stream>>>(k, d_at, (const double*)d_val, (double*)t);
        else
            scatter_kernel<float><<<blocks, 256, 0, H->stream>>>(k, d_at, (const float*)d_val, (float*)t);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(3, hipGetErrorString(e));
    return 0;
}

int w2d_set_active(w2d_handle* H, const int32_t* tiles, int32_t n) {
    if (!H || n < 0 || n > H->ntiles || (n && !tiles)) return fail(1, "set_active: bad arguments");
    if (n == H->nact && std::equal(tiles, tiles + n, H->active_host.begin())) return 0;
    std::vector<char> now(H->ntiles, 0);
    for (int a = 0; a < n; ++a) {
        if (tiles[a] <= 0 || tiles[a] >= H->ntiles) return fail(1, "set_active: tile out of range");
        now[tiles[a]] = 1;
    }
    int s;
    if ((s = run_align(H))) return s;
    // a tile that leaves the list keeps one value in every buffer of a field the reference keeps once
    std::vector<int32_t> leaving;
    for (int t : H->active_host)
        if (!now[t]) leaving.push_back(t);
    if (!leaving.empty()) {
        // the momentum of every listed tile is stored first: a leaving tile's depth buffers are about to be made one
        if ((s = materialize(H))) return s;
        HIP_TRY(hipMemcpyAsync(H->tiles_scratch, leaving.data(), leaving.size() * sizeof(int), hipMemcpyHostToDevice, H->stream));
        if (H->prec == 8)
            leave_tiles_kernel<double><<<(int)leaving.size(), NT, 0, H->stream>>>(
                H->tiles_scratch, (int)leaving.size(), bufs<double>(H), H->par);
        else
            leave_tiles_kernel<float><<<(int)leaving.size(), NT, 0, H->stream>>>(
                H->tiles_scratch, (int)leaving.size(), bufs<float>(H), H->par);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(H->stream));
    }
    H->active_host.assign(tiles, tiles + n);
    H->nact = n;
    if (n) HIP_TRY(hipMemcpyAsync(H->active, tiles, n * sizeof(int), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_set_boundary(w2d_handle* H, const int32_t* cell, const int32_t* ref, const int32_t* kind, const float* d0,
                     const float* h0, int32_t n) {
    if (!H || n < 0 || (n && (!cell || !ref || !kind || !d0 || !h0))) return fail(1, "set_boundary: bad arguments");
    std::vector<int32_t> c(n), r(n);
    for (int m = 0; m < n; ++m) {
        const int64_t a = own_of(H, cell[m]);
        if (a < 0) return fail(1, "set_boundary: a boundary cell is not an own cell");
        if (kind[m] < 0 || kind[m] > 2) return fail(1, "set_boundary: a kind is not 0, 1 or 2");
        // only a following cell (kind 2) reads its reference; the others may carry any index, as the reference's
        // callers pass (a free outflow's ref is 0 there)
        const int64_t b = kind[m] == 2 ? own_of(H, ref[m]) : a;
        if (b < 0) return fail(1, "set_boundary: a following cell's reference is not an own cell");
        c[m] = (int32_t)a;
        r[m] = (int32_t)b;
    }
    int s;
    H->nb_cells = 0;
    if (n) {
        if ((s = upload(H, &H->b_cell, c.data(), n))) return s;
        if ((s = upload(H, &H->b_ref, r.data(), n))) return s;
        if ((s = upload(H, &H->b_kind, kind, n))) return s;
        if ((s = upload(H, &H->b_d0, d0, n))) return s;
        if ((s = upload(H, &H->b_h0, h0, n))) return s;
    }
    H->nb_cells = n;
    H->b_cell_host = c;
    H->b_ref_host = r;
    H->b_kind_host.assign(kind, kind + n);
    return 0;
}

int w2d_step(w2d_handle* H, double dt, int32_t want_wave, int32_t want_wet) {
    (void)want_wet;  // every active tile's flag is written on every step, as the CPU twin does
    if (!H) return fail(1, "step: no handle");
    if (want_wave) {
        if (H->prec == 8)
            fill_kernel<double><<<64, 256, 0, H->stream>>>((double*)H->wave, (size_t)H->ntiles, 0.0);
        else
            fill_kernel<float><<<64, 256, 0, H->stream>>>((float*)H->wave, (size_t)H->ntiles, 0.0f);
        HIP_TRY(hipGetLastError());
    }
    if (H->nact == 0) return 0;
    int s;
    if ((s = run_align(H))) return s;
    if ((s = run_step(H, dt, want_wave))) return s;
    H->par = 1 - H->par;
    H->m_fresh = 0;
    if (H->run_m) H->members_stale = 1;
    return 0;
}

int w2d_max_wave(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "max_wave: bad arguments");
    if (H->nact == 0) {
        *out = 0.0;
        return 0;
    }
    if (H->prec == 8)
        wave_max_kernel<double><<<1, 1024, 0, H->stream>>>(H->active, H->nact, (double*)H->wave, H->scalar);
    else
        wave_max_kernel<float><<<1, 1024, 0, H->stream>>>(H->active, H->nact, (float*)H->wave, H->scalar);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, H->scalar, sizeof(double), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_wave(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "wave: bad arguments");
    if (H->prec == 8) {
        HIP_TRY(hipMemcpyAsync(out, H->wave, H->ntiles * sizeof(double), hipMemcpyDeviceToHost, H->stream));
        HIP_TRY(hipStreamSynchronize(H->stream));
    } else {
        std::vector<float> w(H->ntiles);
        HIP_TRY(hipMemcpyAsync(w.data(), H->wave, H->ntiles * sizeof(float), hipMemcpyDeviceToHost, H->stream));
        HIP_TRY(hipStreamSynchronize(H->stream));
        for (int t = 0; t < H->ntiles; ++t) out[t] = w[t];
    }
    return 0;
}

int w2d_wet_flags(w2d_handle* H, int32_t* out) {
    if (!H || !out) return fail(1, "wet_flags: bad arguments");
    HIP_TRY(hipMemcpyAsync(out, H->wet, H->ntiles * sizeof(int), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_set_wet(w2d_handle* H, const int32_t* flags) {
    if (!H || !flags) return fail(1, "set_wet: bad arguments");
    HIP_TRY(hipMemcpyAsync(H->wet, flags, H->ntiles * sizeof(int), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_refresh_plan(w2d_handle* H, const int32_t* tiles, int32_t n) {
    if (!H || n < 0 || n > H->ntiles || (n && !tiles)) return fail(1, "refresh_plan: bad arguments");
    if (!n) return 0;
    for (int a = 0; a < n; ++a)
        if (tiles[a] <= 0 || tiles[a] >= H->ntiles) return fail(1, "refresh_plan: tile out of range");
    HIP_TRY(hipMemcpyAsync(H->tiles_scratch, tiles, n * sizeof(int), hipMemcpyHostToDevice, H->stream));
    const int last = H->rows - 1;
    int s;
    if ((s = run_align(H))) return s;
    const Cur<char> c = cur(H);
    if (H->prec == 8)
        plan_kernel<double, false><<<n, NT, 0, H->stream>>>(H->tiles_scratch, n, (double*)c.h, (double*)c.om, H->hc,
                                                           H->ctab, H->ncell, last);
    else if (H->fast)
        plan_kernel<float, true><<<n, NT, 0, H->stream>>>(H->tiles_scratch, n, (float*)c.h, (float*)c.om, H->hc,
                                                         H->ctab, H->ncell, last);
    else
        plan_kernel<float, false><<<n, NT, 0, H->stream>>>(H->tiles_scratch, n, (float*)c.h, (float*)c.om, H->hc,
                                                          H->ctab, H->ncell, last);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_refresh_halos(w2d_handle* H) { return H ? 0 : fail(1, "refresh_halos: no handle"); }

int w2d_tally_set(w2d_handle* H, const int32_t* face, const int32_t* axis, const float* sign, int32_t n) {
    if (!H || n < 0 || (n && (!face || !axis || !sign))) return fail(1, "tally_set: bad arguments");
    std::vector<int32_t> f(n);
    for (int m = 0; m < n; ++m) {
        const int64_t c = own_of(H, face[m]);
        if (c < 0) return fail(1, "tally_set: a face is not an own face");
        f[m] = (int32_t)c;
    }
    int s;
    H->n_tally = 0;
    if (n) {
        if ((s = upload(H, &H->t_face, f.data(), n))) return s;
        if ((s = upload(H, &H->t_axis, axis, n))) return s;
        if ((s = upload(H, &H->t_sign, sign, n))) return s;
        if ((s = dalloc(H, (void**)&H->t_total, n * sizeof(double)))) return s;
        if ((s = dalloc(H, (void**)&H->t_flow, n * sizeof(double)))) return s;
        HIP_TRY(hipStreamSynchronize(H->stream));
    }
    H->n_tally = n;
    H->t_face_host = f;
    H->t_axis_host.assign(axis, axis + n);
    return 0;
}

int w2d_tally_add(w2d_handle* H, double dt) {
    if (!H) return fail(1, "tally_add: no handle");
    if (!H->n_tally) return 0;
    const int n = H->n_tally;
    int s;
    if ((s = run_align(H))) return s;
    const Cur<char> c = cur(H);
    if (H->prec == 8)
        tally_kernel<double><<<(n + 255) / 256, 256, 0, H->stream>>>(n, dt, H->t_face, H->t_axis, H->t_sign,
                                                                     (double*)c.qx, (double*)c.qy, H->t_total, H->t_flow);
    else
        tally_kernel<float><<<(n + 255) / 256, 256, 0, H->stream>>>(n, dt, H->t_face, H->t_axis, H->t_sign,
                                                                    (float*)c.qx, (float*)c.qy, H->t_total, H->t_flow);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int sum_doubles(w2d_handle* H, const double* x, int64_t n, double* out) {
    sum_kernel<float><<<1, 1024, 0, H->stream>>>(n, nullptr, nullptr, nullptr, x, H->scalar);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, H->scalar, sizeof(double), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_tally_volume(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "tally_volume: bad arguments");
    if (!H->n_tally) {
        *out = 0.0;
        return 0;
    }
    return sum_doubles(H, H->t_total, H->n_tally, out);
}

int w2d_tally_restore(w2d_handle* H, const double* totals, int32_t n) {
    if (!H || n != H->n_tally || (n && !totals)) return fail(1, "tally_restore: bad arguments (one total per face)");
    if (n) {
        HIP_TRY(hipMemcpyAsync(H->t_total, totals, n * sizeof(double), hipMemcpyHostToDevice, H->stream));
        HIP_TRY(hipStreamSynchronize(H->stream));
    }
    return 0;
}

int w2d_tally_flow(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "tally_flow: bad arguments");
    if (!H->n_tally) {
        *out = 0.0;
        return 0;
    }
    return sum_doubles(H, H->t_flow, H->n_tally, out);
}

int w2d_storage_set(w2d_handle* H, const int64_t* index, int64_t n) {
    if (!H || n < 0 || (n && !index)) return fail(1, "storage_set: bad arguments");
    std::vector<int64_t> own(n);
    for (int64_t m = 0; m < n; ++m) {
        own[m] = own_of(H, index[m]);
        if (own[m] < 0) return fail(1, "storage_set: an index is not an own cell");
    }
    int s;
    H->n_store = 0;
    if (n && (s = upload(H, &H->store_index, own.data(), n))) return s;
    H->n_store = n;
    return 0;
}

int w2d_storage(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "storage: bad arguments");
    if (!H->n_store) {
        *out = 0.0;
        return 0;
    }
    int s;
    if ((s = run_align(H))) return s;
    const Cur<char> c = cur(H);
    if (H->prec == 8) return big_sum<double>(H, H->n_store, H->store_index, (double*)c.h, (double*)c.res, out);
    return big_sum<float>(H, H->n_store, H->store_index, (float*)c.h, (float*)c.res, out);
}

int w2d_maxima_set(w2d_handle* H, double arrive, double deep) {
    if (!H) return fail(1, "maxima_set: no handle");
    H->m_arrive = arrive;
    H->m_deep = deep;
    const size_t es = esize(H);
    int s;
    for (int q = 0; q < 5; ++q) {
        if (!H->maxima[q] && (s = dalloc(H, &H->maxima[q], H->ncell * es))) return s;
        if (H->prec == 8)
            fill_kernel<double><<<256, 256, 0, H->stream>>>((double*)H->maxima[q], H->ncell, q == 4 ? -1.0 : 0.0);
        else
            fill_kernel<float><<<256, 256, 0, H->stream>>>((float*)H->maxima[q], H->ncell, q == 4 ? -1.0f : 0.0f);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_maxima_update(w2d_handle* H, double t) {
    if (!H || !H->maxima[0]) return fail(1, "maxima_update: call maxima_set first");
    int s;
    if ((s = run_align(H))) return s;
    if (!H->nact) return 0;
    const Cur<char> c = cur(H);
    if (H->prec == 8)
        maxima_kernel<double><<<H->nact, NT, 0, H->stream>>>(
            H->active, H->nact, H->nbr, t, H->m_arrive, H->m_deep, (double*)c.d, (double*)c.h, (double*)c.u,
            (double*)c.v, (double*)H->maxima[0], (double*)H->maxima[1], (double*)H->maxima[2], (double*)H->maxima[3],
            (double*)H->maxima[4]);
    else
        maxima_kernel<float><<<H->nact, NT, 0, H->stream>>>(
            H->active, H->nact, H->nbr, (float)t, (float)H->m_arrive, (float)H->m_deep, (float*)c.d, (float*)c.h,
            (float*)c.u, (float*)c.v, (float*)H->maxima[0], (float*)H->maxima[1], (float*)H->maxima[2],
            (float*)H->maxima[3], (float*)H->maxima[4]);
    HIP_TRY(hipGetLastError());
    return 0;
}

int w2d_maxima_get(w2d_handle* H, int32_t which, void* host, int64_t n) {
    if (!H || which < 0 || which > 4 || !H->maxima[which] || n != (int64_t)H->ntiles * STRIDE2)
        return fail(1, "maxima_get: bad arguments");
    const size_t es = esize(H);
    std::vector<char> own(H->ncell * es);
    HIP_TRY(hipMemcpyAsync(own.data(), H->maxima[which], own.size(), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    if (es == 8)
        expand(H, (const double*)own.data(), (double*)host);
    else
        expand(H, (const float*)own.data(), (float*)host);
    return 0;
}

int w2d_maxima_put(w2d_handle* H, int32_t which, const void* host, int64_t n) {
    if (!H || !host || which < 0 || which > 4 || !H->maxima[which] || n != (int64_t)H->ntiles * STRIDE2)
        return fail(1, "maxima_put: bad arguments (call maxima_set first)");
    const size_t es = esize(H);
    std::vector<char> own(H->ncell * es);
    if (es == 8)
        compact(H, (const double*)host, (double*)own.data());
    else
        compact(H, (const float*)host, (float*)own.data());
    HIP_TRY(hipMemcpyAsync(H->maxima[which], own.data(), own.size(), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_source_weights(w2d_handle* H, const int64_t* index, const double* w, int64_t n) {
    if (!H || n < 0 || (n && (!index || !w))) return fail(1, "source_weights: bad arguments");
    std::vector<double> own(H->ncell, 0.0);
    for (int64_t m = 0; m < n; ++m) {
        const int64_t c = own_of(H, index[m]);
        if (c < 0) return fail(1, "source_weights: an index is not an own cell");
        own[c] = w[m];
    }
    H->weight_host = own;
    H->excess = 0;
    if (!H->src_weight) {
        int s = dalloc(H, (void**)&H->src_weight, H->ncell * sizeof(double));
        if (s) return s;
    }
    HIP_TRY(hipMemcpyAsync(H->src_weight, own.data(), H->ncell * sizeof(double), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_source_scale(w2d_handle* H, double scale) {
    if (!H || !H->src_weight) return fail(1, "source_scale: call source_weights first");
    if (H->excess) return fail(1, "source_scale: the source is in the excess form (w2d_source_losses)");
    if (H->prec == 8)
        scale_source_kernel<double><<<1024, 256, 0, H->stream>>>(H->ncell, H->src_weight, scale, (double*)H->src);
    else
        scale_source_kernel<float><<<1024, 256, 0, H->stream>>>(H->ncell, H->src_weight, scale, (float*)H->src);
    HIP_TRY(hipGetLastError());
    return 0;
}

int w2d_source_sum(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "source_sum: bad arguments");
    if (H->prec == 8) return big_sum<double>(H, (int64_t)H->ncell, nullptr, (double*)H->src, nullptr, out);
    return big_sum<float>(H, (int64_t)H->ncell, nullptr, (float*)H->src, nullptr, out);
}

int w2d_source_losses(w2d_handle* H, const int64_t* index, const double* w, const double* cn, int64_t n) {
    if (!H || n < 0 || (n && (!index || !w || !cn))) return fail(1, "source_losses: bad arguments");
    std::vector<double> own_w(H->ncell, 0.0), own_cn(H->ncell, std::nan(""));
    for (int64_t m = 0; m < n; ++m) {
        const int64_t c = own_of(H, index[m]);
        if (c < 0) return fail(1, "source_losses: an index is not an own cell");
        if (!std::isfinite(w[m])) return fail(1, "source_losses: a depth factor is not finite");
        if (!std::isnan(cn[m]) && !(cn[m] > 0.0 && cn[m] <= 100.0))
            return fail(1, "source_losses: a curve number is outside (0, 100]");
        own_w[c] = w[m];
        own_cn[c] = cn[m];
    }
    int s;
    if (!H->src_weight && (s = dalloc(H, (void**)&H->src_weight, H->ncell * sizeof(double)))) return s;
    if (!H->src_cn && (s = dalloc(H, (void**)&H->src_cn, H->ncell * sizeof(double)))) return s;
    const size_t nt = (size_t)H->ntiles;
    if (!H->ex_partial && (s = dalloc(H, (void**)&H->ex_partial, nt * sizeof(double)))) return s;
    if (!H->ex_flag && (s = dalloc(H, (void**)&H->ex_flag, nt * sizeof(int)))) return s;
    HIP_TRY(hipMemcpyAsync(H->src_weight, own_w.data(), H->ncell * sizeof(double), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipMemcpyAsync(H->src_cn, own_cn.data(), H->ncell * sizeof(double), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    H->weight_host = own_w;
    H->excess = 1;
    if (H->src_pix) {  // a new excess form starts without pixels: w2d_source_pixels sets them again
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->src_pix), H->owned.end());
        (void)hipFree(H->src_pix);
        H->src_pix = nullptr;
        H->npix = 0;
    }
    return 0;
}

}  // extern "C"

namespace {

int source_excess_any(w2d_handle* H, double d0, double d1, const int* pix, const double* pdep, int npix, double denom,
                      double* total, int32_t* tiles) {
    const int nt = H->ntiles;
    if (H->prec == 8)
        excess_kernel<double><<<std::min(nt, 2048), EX_THREADS, 0, H->stream>>>(
            nt, nullptr, 0, nullptr, nullptr, d0, d1, denom, H->src_weight, H->src_cn, pix, pdep, npix,
            (double*)H->src, H->ex_partial, H->ex_flag);
    else
        excess_kernel<float><<<std::min(nt, 2048), EX_THREADS, 0, H->stream>>>(
            nt, nullptr, 0, nullptr, nullptr, d0, d1, denom, H->src_weight, H->src_cn, pix, pdep, npix,
            (float*)H->src, H->ex_partial, H->ex_flag);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(nt, 0.0);
    std::vector<int32_t> flag(nt, 0);
    HIP_TRY(hipMemcpyAsync(part.data(), H->ex_partial, nt * sizeof(double), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipMemcpyAsync(flag.data(), H->ex_flag, nt * sizeof(int), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    double sum = 0.0;
    for (int t = 1; t < nt; ++t) sum += part[t];
    *total = sum;
    if (tiles) {
        tiles[0] = 0;
        for (int t = 1; t < nt; ++t) tiles[t] = flag[t];
    }
    return 0;
}

}  // namespace

extern "C" {

int w2d_source_excess(w2d_handle* H, double d0, double d1, double denom, double* total, int32_t* tiles) {
    if (!H || !total || !(denom > 0.0)) return fail(1, "source_excess: bad arguments");
    if (!H->excess) return fail(1, "source_excess: call source_losses first");
    if (H->src_pix) return fail(1, "source_excess: the source is gridded (w2d_source_excess_pixels)");
    return source_excess_any(H, d0, d1, nullptr, nullptr, 0, denom, total, tiles);
}

int w2d_source_pixels(w2d_handle* H, const int64_t* index, const int32_t* pixel, int64_t n, int32_t npix) {
    if (!H || n < 0 || npix < 1 || (n && (!index || !pixel))) return fail(1, "source_pixels: bad arguments");
    if (!H->excess) return fail(1, "source_pixels: call source_losses first");
    std::vector<int32_t> own(H->ncell, -1);
    for (int64_t m = 0; m < n; ++m) {
        const int64_t c = own_of(H, index[m]);
        if (c < 0) return fail(1, "source_pixels: an index is not an own cell");
        if (pixel[m] < 0 || pixel[m] >= npix) return fail(1, "source_pixels: a pixel is outside [0, npix)");
        own[c] = pixel[m];
    }
    int s;
    if (!H->src_pix && (s = dalloc(H, (void**)&H->src_pix, H->ncell * sizeof(int32_t)))) return s;
    HIP_TRY(hipMemcpyAsync(H->src_pix, own.data(), H->ncell * sizeof(int32_t), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    H->npix = npix;
    return 0;
}

int w2d_source_excess_pixels(w2d_handle* H, const double* d0, const double* d1, int32_t npix, double denom,
                             double* total, int32_t* tiles) {
    if (!H || !d0 || !d1 || !total || !(denom > 0.0)) return fail(1, "source_excess_pixels: bad arguments");
    if (!H->excess || !H->src_pix) return fail(1, "source_excess_pixels: call source_losses and source_pixels first");
    if (npix != H->npix) return fail(1, "source_excess_pixels: npix is not source_pixels' count");
    int s;
    if (H->pix_dep_n < 2 * (size_t)npix) {
        if (H->pix_dep) {
            H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->pix_dep), H->owned.end());
            (void)hipFree(H->pix_dep);
            H->pix_dep = nullptr;
        }
        if ((s = dalloc(H, (void**)&H->pix_dep, 2 * (size_t)npix * sizeof(double)))) return s;
        H->pix_dep_n = 2 * (size_t)npix;
    }
    HIP_TRY(hipMemcpyAsync(H->pix_dep, d0, npix * sizeof(double), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipMemcpyAsync(H->pix_dep + npix, d1, npix * sizeof(double), hipMemcpyHostToDevice, H->stream));
    return source_excess_any(H, 0.0, 0.0, H->src_pix, H->pix_dep, npix, denom, total, tiles);
}

int w2d_run_pixel_terms(w2d_handle* H, const double* depths, int64_t nwindows, int32_t npix) {
    if (!H || nwindows < 0 || (nwindows && !depths)) return fail(1, "run_pixel_terms: bad arguments");
    if (!H->src_pix || npix != H->npix) return fail(1, "run_pixel_terms: call source_pixels first, with this npix");
    H->run_pix_host.assign(depths, depths + 2 * (size_t)npix * nwindows);
    return 0;
}

int w2d_run_terms(w2d_handle* H, const double* terms, int64_t nwindows) {
    if (!H || nwindows < 0 || (nwindows && !terms)) return fail(1, "run_terms: bad arguments");
    for (int64_t k = 0; k < nwindows; ++k)
        if (!(terms[3 * k + 2] > 0.0)) return fail(1, "run_terms: a window's 1000 (t1 - t0) is not positive");
    H->run_terms_host.assign(terms, terms + 3 * nwindows);
    return 0;
}

int w2d_timing(w2d_handle* H, int32_t on) {
    if (!H) return fail(1, "timing: no handle");
    H->timing = on != 0;
    return 0;
}

static int timing_collect(w2d_handle* H, double* faces_ms, double* cells_ms, int64_t* launches);

int w2d_timing_read(w2d_handle* H, double* faces_ms, double* cells_ms, int64_t* launches) {
    if (!H || !faces_ms || !cells_ms || !launches) return fail(1, "timing_read: bad arguments");
    return timing_collect(H, faces_ms, cells_ms, launches);
}

// The summed milliseconds between the recorded marks (three per step), cleared; the caller checked the arguments.
static int timing_collect(w2d_handle* H, double* faces_ms, double* cells_ms, int64_t* launches) {
    HIP_TRY(hipStreamSynchronize(H->stream));
    double f = 0.0, c = 0.0;
    const size_t n = H->marks.size() / 3;
    for (size_t q = 0; q < n; ++q) {
        float a = 0.0f, b = 0.0f;
        HIP_TRY(hipEventElapsedTime(&a, H->marks[3 * q], H->marks[3 * q + 1]));
        HIP_TRY(hipEventElapsedTime(&b, H->marks[3 * q + 1], H->marks[3 * q + 2]));
        f += a;
        c += b;
    }
    for (hipEvent_t e : H->marks) H->event_pool.push_back(e);
    H->marks.clear();
    *faces_ms = f;
    *cells_ms = c;
    *launches = (int64_t)n;
    return 0;
}

int w2d_copy_probe(int64_t bytes, int32_t repeats, double* ms) {
    if (bytes < 16 || repeats < 1 || !ms) return fail(1, "copy_probe: bad arguments");
    const size_t n = (size_t)bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIP_TRY(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(2, "copy_probe: out of memory");
    }
    HIP_TRY(hipMemset(a, 0, n * 16));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    copy_probe_kernel<<<blocks, 256>>>(n, a, b);
    HIP_TRY(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; ++r) copy_probe_kernel<<<blocks, 256>>>(n, a, b);
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

int w2d_closure_set(w2d_handle* H, int32_t n, const int32_t* cell_r, const int32_t* cell_l, const int32_t* axis,
                    const float* stat, const float* norm, const float* params, int32_t conveyance) {
    if (!H) return fail(1, "closure_set: no handle");
    if (n < 0 || (n && (!cell_r || !cell_l || !axis || !stat || !norm || !params)))
        return fail(1, "closure_set: bad arguments");
    if (n && H->prec != 4) return fail(1, "closure_set: the learned closure runs in FP32");
    for (int f = 0; f < n; ++f)
        if (cell_r[f] < (int)TILE2 || (size_t)cell_r[f] >= H->ncell || cell_l[f] < (int)TILE2 ||
            (size_t)cell_l[f] >= H->ncell || (axis[f] != 0 && axis[f] != 1))
            return fail(1, "closure_set: a face is out of range");
    auto drop = [&](void* p) {
        if (!p) return;
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), p), H->owned.end());
        (void)hipFree(p);
    };
    drop(H->cl_cell_r);
    drop(H->cl_cell_l);
    drop(H->cl_axis);
    drop(H->cl_stat);
    drop(H->cl_norm);
    drop(H->cl_params);
    drop(H->cl_wbf);
    H->cl_cell_r = H->cl_cell_l = H->cl_axis = nullptr;
    H->cl_stat = H->cl_norm = H->cl_params = nullptr;
    H->cl_wbf = nullptr;
    if (H->graph) {  // a captured graph holds the old launches
        (void)hipGraphExecDestroy(H->graph);
        H->graph = nullptr;
        H->graph_steps = 0;
    }
    H->closure_n = 0;
    if (n == 0) return 0;
    int s;
    using namespace w2d_mlp;
    std::vector<uint16_t> wbf(NBF16, 0);
    for (int e = 0; e < HID * IN; ++e) wbf[BF_W1 + e] = to_bf16(params[OFF_W1 + e]);
    for (int e = 0; e < HID * HID; ++e) wbf[BF_W2 + e] = to_bf16(params[OFF_W2 + e]);
    for (int e = 0; e < OUT * HID; ++e) wbf[BF_W3 + e] = to_bf16(params[OFF_W3 + e]);
    if ((s = upload(H, &H->cl_cell_r, cell_r, n)) || (s = upload(H, &H->cl_cell_l, cell_l, n)) ||
        (s = upload(H, &H->cl_axis, axis, n)) || (s = upload(H, &H->cl_stat, stat, (size_t)n * CLOSURE_STATIC)) ||
        (s = upload(H, &H->cl_norm, norm, 2 * IN + 1)) || (s = upload(H, &H->cl_params, params, NPARAM)) ||
        (s = upload(H, &H->cl_wbf, wbf.data(), NBF16)))
        return s;
    // every face starts uncorrected: a face left out of this set keeps no correction from an earlier one
    std::vector<float2> one(H->ncell, make_float2(1.f, 0.f));
    if (!H->corr_x) {
        if ((s = upload(H, &H->corr_x, one.data(), one.size())) || (s = upload(H, &H->corr_y, one.data(), one.size())))
            return s;
    } else {
        HIP_TRY(hipMemcpy(H->corr_x, one.data(), one.size() * sizeof(float2), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(H->corr_y, one.data(), one.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    H->closure_n = n;
    H->closure_conveyance = conveyance != 0;
    return 0;
}

int w2d_set_launch(w2d_handle* H, int32_t faces_threads, int32_t cells_threads) {
    if (!H) return fail(1, "set_launch: no handle");
    if (faces_threads != 256 && faces_threads != 512) return fail(1, "set_launch: faces_threads is 256 or 512");
    if (cells_threads != 256 && cells_threads != 512) return fail(1, "set_launch: cells_threads is 256 or 512");
    if ((faces_threads != H->faces_threads || cells_threads != H->cells_threads) && H->graph) {
        // a captured graph holds the old launches
        (void)hipGraphExecDestroy(H->graph);
        H->graph = nullptr;
        H->graph_steps = 0;
    }
    H->faces_threads = faces_threads;
    H->cells_threads = cells_threads;
    return 0;
}

int w2d_synchronize(w2d_handle* H) {
    if (!H) return fail(1, "synchronize: no handle");
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

}  // extern "C"

// ---- device-resident run: host side


namespace {

RunLists run_lists(w2d_handle* H) {
    return RunLists{H->nbr, H->wet, H->run_has_source, H->run_bnd_tile, H->run_is_active, H->run_idle,
                    H->run_flags, H->run_seg, H->run_woken, H->run_leaving, H->active};
}

template <class R, bool FAST>
int enqueue_pre(w2d_handle* H, int flush_only) {
    pre_kernel<R, FAST><<<H->run_m, PRE_THREADS, 0, H->stream>>>(
        H->run_mem, H->run_glob, H->run_scales, H->run_num, run_lists(H), bufs<R>(H), H->hc, H->ctab, H->ncell,
        H->rows - 1, (R)H->P.dx, H->m_fresh, H->run_wave, H->run_src_partial, flush_only);
    HIP_TRY(hipGetLastError());
    return 0;
}

PipeArgs pipe_args(w2d_handle* H, int drain) {
    return PipeArgs{H->run_is_active, H->pipe_near, H->pipe_bidx, H->pipe_tidx, H->b_kind, H->b_ref, H->b_d0,
                    H->b_h0, H->t_sign, H->t_total, H->t_flow, drain};
}

template <class R>
StepArgs<R> run_step_args(w2d_handle* H) {
    StepArgs<R> A = step_args<R>(H, 0.0, 0);
    A.members = H->run_mem;
    A.tile_member = H->run_tile_member;
    A.glob = H->run_glob;
    A.src_weight = H->src_weight;
    A.src_partial = H->run_src_partial;
    A.member_wave = H->run_wave;
    A.excess = H->excess;
    if (H->run_maxima) {
        A.every = H->run_maxima;
        A.arrive = (R)H->m_arrive;
        A.deep = (R)H->m_deep;
        A.mx_d = (R*)H->maxima[0];
        A.mx_h = (R*)H->maxima[1];
        A.mx_s = (R*)H->maxima[2];
        A.mx_z = (R*)H->maxima[3];
        A.mx_t = (R*)H->maxima[4];
        A.bnd_tile = H->run_bnd_tile;
    }
    return A;
}

// The pipe kernel in place of the face kernel (a run whose members pipeline), then the cell kernel for the members
// whose update this step is not left to the next step.
template <class R, bool FAST>
int launch_pipe_kernels(w2d_handle* H, const StepArgs<R>& A, int blocks) {
    int s;
    if (H->timing && (s = mark(H))) return s;
    pipe_kernel<R, FAST, 512><<<blocks, 512, 0, H->stream>>>(A, pipe_args(H, 0));
    HIP_TRY(hipGetLastError());
    if (H->timing && (s = mark(H))) return s;
    if (H->cells_threads == 512)
        cells_kernel<R, FAST, 512><<<blocks, 512, 0, H->stream>>>(A);
    else
        cells_kernel<R, FAST, 256><<<blocks, 256, 0, H->stream>>>(A);
    HIP_TRY(hipGetLastError());
    if (H->timing && (s = mark(H))) return s;
    return 0;
}

// The updates still waiting after the last step (pipe kernel with no face step) and their books (post kernel).
template <class R, bool FAST>
int enqueue_drain(w2d_handle* H) {
    const StepArgs<R> A = run_step_args<R>(H);
    pipe_kernel<R, FAST, 512><<<H->ntiles, 512, 0, H->stream>>>(A, pipe_args(H, 1));
    HIP_TRY(hipGetLastError());
    post_kernel<R><<<H->run_m, POST_THREADS, 0, H->stream>>>(H->run_mem, H->run_wave, H->run_is_active,
                                                     H->run_src_partial, H->run_num, H->b_cell, H->b_ref, H->b_kind,
                                                     H->b_d0, H->b_h0, A.b, H->t_face, H->t_axis, H->t_sign,
                                                     H->t_total, H->t_flow, 1, H->run_glob);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class R, bool FAST>
int enqueue_run_step(w2d_handle* H) {
    const StepArgs<R> A = run_step_args<R>(H);
    const int nt = H->ntiles;
    int s;
    if (H->excess) {
        excess_kernel<R><<<std::min(nt, 2048), EX_THREADS, 0, H->stream>>>(
            nt, H->run_mem, H->run_m, H->run_tile_member, H->run_terms, 0.0, 0.0, 1.0, H->src_weight, H->src_cn,
            H->src_pix, H->run_pix_terms, H->npix, (R*)H->src, H->run_src_partial, H->run_has_source);
        HIP_TRY(hipGetLastError());
    }
    if ((s = enqueue_pre<R, FAST>(H, 0))) return s;
    if ((s = H->run_pipe ? launch_pipe_kernels<R, FAST>(H, A, nt) : launch_step_kernels<R, FAST>(H, A, nt))) return s;
    post_kernel<R><<<H->run_m, POST_THREADS, 0, H->stream>>>(H->run_mem, H->run_wave, H->run_is_active,
                                                     H->run_src_partial, H->run_num, H->b_cell, H->b_ref, H->b_kind,
                                                     H->b_d0, H->b_h0, A.b, H->t_face, H->t_axis, H->t_sign,
                                                     H->t_total, H->t_flow, 0, H->run_glob);
    if (H->run_maxima && H->run_nbnd)
        maxima_run_kernel<R><<<H->run_nbnd, NT, 0, H->stream>>>(
            H->run_bnd_list, H->run_is_active, H->run_mem, H->run_tile_member, H->nbr, (R)H->m_arrive,
            (R)H->m_deep, A.b, (R*)H->maxima[0], (R*)H->maxima[1], (R*)H->maxima[2], (R*)H->maxima[3],
            (R*)H->maxima[4], H->run_maxima);
    HIP_TRY(hipGetLastError());
    H->m_fresh = 0;
    H->run_dirty = 1;
    return 0;
}

int run_flush(w2d_handle* H) {
    if (H->prec == 8) return enqueue_pre<double, false>(H, 1);
    if (H->fast) return enqueue_pre<float, true>(H, 1);
    return enqueue_pre<float, false>(H, 1);
}

int run_step_any(w2d_handle* H) {
    if (H->prec == 8) return enqueue_run_step<double, false>(H);
    if (H->fast) return enqueue_run_step<float, true>(H);
    return enqueue_run_step<float, false>(H);
}

int run_drain(w2d_handle* H) {
    if (!H->run_pipe) return 0;
    if (H->prec == 8) return enqueue_drain<double, false>(H);
    if (H->fast) return enqueue_drain<float, true>(H);
    return enqueue_drain<float, false>(H);
}

// The members as the device holds them, brought to the handle's parities where asked (after the handle stepped on
// its own, its buffers are the current ones).
int members_read(w2d_handle* H, std::vector<Member>& mem) {
    mem.resize(H->run_m);
    HIP_TRY(hipMemcpyAsync(mem.data(), H->run_mem, H->run_m * sizeof(Member), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}
int members_write(w2d_handle* H, const std::vector<Member>& mem) {
    HIP_TRY(hipMemcpyAsync(H->run_mem, mem.data(), H->run_m * sizeof(Member), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int run_align(w2d_handle* H) {
    if (!H->run_m || !H->run_dirty) return 0;
    int s;
    if ((s = run_flush(H))) return s;  // a check the last step asked for, as the host loop makes after its step
    std::vector<Member> mem;
    if ((s = members_read(H, mem))) return s;
    RunGlobals g;
    HIP_TRY(hipMemcpy(&g, H->run_glob, sizeof(g), hipMemcpyDeviceToHost));
    const int par = mem[0].par;
    bool moved = false;
    for (int m = 0; m < H->run_m; ++m) {
        Member& b = mem[m];
        if (b.par == par) continue;
        const int cells = (b.tile1 - b.tile0) * TILE2;
        const int blocks = std::min((cells + 255) / 256, 4096);
        if (H->prec == 8)
            align_kernel<double><<<blocks, 256, 0, H->stream>>>(b.tile0, b.tile1, bufs<double>(H), b.par, par);
        else
            align_kernel<float><<<blocks, 256, 0, H->stream>>>(b.tile0, b.tile1, bufs<float>(H), b.par, par);
        HIP_TRY(hipGetLastError());
        b.par = par;
        moved = true;
    }
    if (moved && (s = members_write(H, mem))) return s;
    H->par = par;
    H->run_dirty = 0;
    H->members_stale = 0;
    // the protocol's view of the list follows the device's
    H->nact = g.nact;
    H->active_host.resize(g.nact);
    if (g.nact) HIP_TRY(hipMemcpy(H->active_host.data(), H->active, g.nact * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

// Members take the handle's parities when it stepped on its own since they last did (their buffers are the handle's).
int members_follow(w2d_handle* H) {
    if (!H->members_stale) return 0;
    std::vector<Member> mem;
    int s;
    if ((s = members_read(H, mem))) return s;
    for (Member& b : mem) {
        b.par = H->par;
    }
    if ((s = members_write(H, mem))) return s;
    H->members_stale = 0;
    return 0;
}

}  // namespace

extern "C" {

int w2d_run_setup(w2d_handle* H, const w2d_member* members, int32_t M, const double* scales, int64_t nscales,
                  const w2d_run_numerics* num, const int32_t* idle) {
    if (!H || !members || M < 1 || !num || (nscales && !scales) || num->dt_every < 1 || num->active_every < 1)
        return fail(1, "run_setup: bad arguments");
    const int nt = H->ntiles;
    int s;
    if ((s = run_align(H))) return s;  // a run set up before ends where it stands
    std::vector<int32_t> tm(nt, 0);
    std::vector<Member> mem(M);
    for (int m = 0; m < M; ++m) {
        const w2d_member& d = members[m];
        if (d.tile0 < 1 || d.tile1 > nt || d.tile0 >= d.tile1 || d.window <= 0.0 || d.nwindows < 0 ||
            d.scale_offset < 0 || d.scale_offset + d.nwindows > nscales)
            return fail(1, "run_setup: a member's tiles or windows are out of range");
        if (m > 0 && d.tile0 < members[m - 1].tile1) return fail(1, "run_setup: members overlap or are out of order");
        for (int t = d.tile0; t < d.tile1; ++t) tm[t] = m;
        Member& b = mem[m];
        std::memset(&b, 0, sizeof(b));
        b.t = d.t0;
        b.t_end = d.t_end;
        b.dt = d.dt0;
        b.window_start = d.t0;
        b.window_end = d.t0;
        b.window_len = d.window;
        b.nsteps = d.nsteps0;
        b.scale_off = d.scale_offset;
        b.nwin = d.nwindows;
        b.tile0 = d.tile0;
        b.tile1 = d.tile1;
        b.refresh_due = d.refresh_due;
        b.dry = d.dry != 0;
        b.dt_min = INFINITY;
        b.t_stop = INFINITY;
        b.par = H->par;
    }
    // tiles that keep themselves active: a source, a boundary cell (model.py:117-122, 145-150); each member's
    // boundary cells and tally faces must be one run of the lists, in member order
    std::vector<int32_t> has_source(nt, 0), bnd_tile(nt, 0), is_active(nt, 0);
    if (H->excess && H->run_terms_host.size() != 3 * (size_t)nscales)
        return fail(1, "run_setup: the excess source needs w2d_run_terms for every window first");
    if (!H->excess && !H->weight_host.empty())
        for (size_t c = 0; c < H->ncell; ++c)
            if (H->weight_host[c] != 0.0) has_source[c / TILE2] = 1;
    auto ranges = [&](const int* dev, int n, bool cells, const char* what) -> int {
        std::vector<int32_t> v(std::max(n, 1));
        if (n) HIP_TRY(hipMemcpy(v.data(), dev, n * sizeof(int), hipMemcpyDeviceToHost));
        int q = 0;
        for (int m = 0; m < M; ++m) {
            const int first = q;
            while (q < n && tm[v[q] / TILE2] == m && v[q] / TILE2 >= members[m].tile0 && v[q] / TILE2 < members[m].tile1) {
                if (cells) bnd_tile[v[q] / TILE2] = 1;
                ++q;
            }
            if (cells) {
                mem[m].b0 = first;
                mem[m].b1 = q;
            } else {
                mem[m].f0 = first;
                mem[m].f1 = q;
            }
        }
        if (q != n) return fail(1, std::string("run_setup: the ") + what + " are not grouped by member in member order");
        return 0;
    };
    if ((s = ranges(H->b_cell, H->nb_cells, true, "boundary cells"))) return s;
    if ((s = ranges(H->t_face, H->n_tally, false, "tally faces"))) return s;
    for (int m = 0; m < M; ++m)
        for (int t = members[m].tile0; t < members[m].tile1; ++t) {
            mem[m].kept_rain += has_source[t] || bnd_tile[t];
            mem[m].kept_dry += bnd_tile[t];
        }
    // each member's neighbours are its own tiles or the void tile: members never touch each other's tiles
    for (int m = 0; m < M; ++m)
        for (int t = members[m].tile0; t < members[m].tile1; ++t)
            for (int q = 0; q < 8; ++q) {
                const int u = H->nbr_host[(size_t)t * 8 + q];
                if (u != 0 && (u < members[m].tile0 || u >= members[m].tile1))
                    return fail(1, "run_setup: a member's tile has a neighbour outside the member");
            }
    std::vector<int32_t> seg(nt, 0);
    for (int t : H->active_host) {
        int m = 0;
        while (m < M && !(t >= members[m].tile0 && t < members[m].tile1)) ++m;
        if (m == M) return fail(1, "run_setup: an active tile belongs to no member");
        is_active[t] = 1;
        seg[members[m].tile0 + mem[m].nlist++] = t;
    }
    std::vector<int32_t> idle_v(nt, 0);
    if (idle) idle_v.assign(idle, idle + nt);
    H->run_m = 0;
    auto up = [&](int** dst, const std::vector<int32_t>& v) -> int {
        if (*dst) {
            HIP_TRY(hipMemcpy(*dst, v.data(), std::max<size_t>(v.size(), 1) * sizeof(int32_t), hipMemcpyHostToDevice));
            return 0;
        }
        return upload(H, dst, v.data(), std::max<size_t>(v.size(), 1));
    };
    std::vector<int32_t> zeros(nt, 0);
    if ((s = up(&H->run_tile_member, tm)) || (s = up(&H->run_has_source, has_source)) ||
        (s = up(&H->run_bnd_tile, bnd_tile)) || (s = up(&H->run_is_active, is_active)) ||
        (s = up(&H->run_idle, idle_v)) || (s = up(&H->run_flags, zeros)) || (s = up(&H->run_woken, zeros)) ||
        (s = up(&H->run_leaving, zeros)) || (s = up(&H->run_seg, seg)))
        return s;
    std::vector<int32_t> bnd_list;
    for (int t = 0; t < nt; ++t)
        if (bnd_tile[t]) bnd_list.push_back(t);
    if (H->run_bnd_list) {
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_bnd_list), H->owned.end());
        (void)hipFree(H->run_bnd_list);
        H->run_bnd_list = nullptr;
    }
    H->run_nbnd = (int)bnd_list.size();
    if (bnd_list.empty()) bnd_list.push_back(0);
    if ((s = upload(H, &H->run_bnd_list, bnd_list.data(), bnd_list.size()))) return s;
    // The pipelined step for every member whose boundary cells allow it: each tile recomputes the update of its west
    // column and north row, so a following cell (kind 2) there must find its reference cell among the cells the tile
    // updates (its own, its west neighbour's east column, its north neighbour's south row), and a reference is never a
    // boundary cell; the learned closure reads the levels before the face step, so it runs the two-kernel step.
    H->run_pipe = 0;
    H->pipe_ok.assign(M, 0);
    if (H->closure_n == 0) {  // the indices are set up whether asked or not: w2d_run_pipeline may switch mid-run
        const size_t n = H->ncell;
        std::vector<int32_t> bidx(n, -1), near(nt, 0);
        std::vector<int2> tidx(n, make_int2(-1, -1));
        for (int p = 0; p < H->nb_cells; ++p) bidx[H->b_cell_host[p]] = p;
        auto nbr_at = [&](int t, int q) { return H->nbr_host[(size_t)t * 8 + q]; };  // NEIGHBOURS order
        auto updated_by = [&](int t, int c) {  // whether tile t updates cell c
            const int tc = c / TILE2, l = c - tc * TILE2;
            return tc == t || (tc != 0 && tc == nbr_at(t, 3) && (l & (TILE - 1)) == TILE - 1) ||
                   (tc != 0 && tc == nbr_at(t, 1) && l / TILE == TILE - 1);
        };
        std::vector<int32_t> ok(M, 1);
        for (int p = 0; p < H->nb_cells; ++p) {
            const int c = H->b_cell_host[p], tc = c / TILE2, l = c - tc * TILE2;
            near[tc] |= 1;
            if (nbr_at(tc, 4)) near[nbr_at(tc, 4)] |= 1;  // east: the cell is in its west column when l % 32 == 31
            if (nbr_at(tc, 6)) near[nbr_at(tc, 6)] |= 1;  // south
            if (H->b_kind_host[p] != 2) continue;
            const int rf = H->b_ref_host[p];
            bool fine = bidx[rf] < 0 && updated_by(tc, rf);
            if ((l & (TILE - 1)) == TILE - 1 && nbr_at(tc, 4)) fine = fine && updated_by(nbr_at(tc, 4), rf);
            if (l / TILE == TILE - 1 && nbr_at(tc, 6)) fine = fine && updated_by(nbr_at(tc, 6), rf);
            if (!fine) ok[tm[tc]] = 0;
        }
        for (int q = 0; q < H->n_tally; ++q) {
            const int c = H->t_face_host[q];
            near[c / TILE2] |= 2;
            if (H->t_axis_host[q])
                tidx[c].y = q;
            else
                tidx[c].x = q;
        }
        H->pipe_ok = ok;
        for (int m = 0; m < M; ++m) {
            mem[m].pipe = H->pipe_on && ok[m];
            H->run_pipe |= mem[m].pipe;
        }
        {
            if (!H->pipe_near && (s = dalloc(H, (void**)&H->pipe_near, (size_t)nt * sizeof(int)))) return s;
            if (!H->pipe_bidx && (s = dalloc(H, (void**)&H->pipe_bidx, n * sizeof(int)))) return s;
            if (!H->pipe_tidx && (s = dalloc(H, (void**)&H->pipe_tidx, n * sizeof(int2)))) return s;
            HIP_TRY(hipMemcpy(H->pipe_near, near.data(), (size_t)nt * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(H->pipe_bidx, bidx.data(), n * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(H->pipe_tidx, tidx.data(), n * sizeof(int2), hipMemcpyHostToDevice));
        }
    }
    if (H->run_mem) {
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_mem), H->owned.end());
        (void)hipFree(H->run_mem);
        H->run_mem = nullptr;
    }
    if ((s = upload(H, &H->run_mem, mem.data(), M))) return s;
    if (H->run_scales) {
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_scales), H->owned.end());
        (void)hipFree(H->run_scales);
        H->run_scales = nullptr;
    }
    std::vector<double> sc(scales, scales + nscales);
    if (sc.empty()) sc.push_back(0.0);
    if ((s = upload(H, &H->run_scales, sc.data(), sc.size()))) return s;
    if (!H->run_glob) {
        RunGlobals g0{};
        if ((s = upload(H, &H->run_glob, &g0, 1))) return s;
    }
    RunGlobals g{};
    g.nact = H->nact;
    HIP_TRY(hipMemcpy(H->run_glob, &g, sizeof(g), hipMemcpyHostToDevice));
    if (!H->run_src_partial && (s = dalloc(H, (void**)&H->run_src_partial, (size_t)nt * sizeof(double)))) return s;
    if (H->run_wave) {
        H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_wave), H->owned.end());
        (void)hipFree(H->run_wave);
        H->run_wave = nullptr;
    }
    if ((s = dalloc(H, (void**)&H->run_wave, (size_t)M * sizeof(unsigned long long)))) return s;
    H->run_num = RunNum{num->cfl, num->dt_max, num->dt_min, num->wave_floor, H->P.dx, H->P.dx * H->P.dx,
                        num->dt_every, num->active_every, H->excess};
    if (H->excess) {
        if (H->run_terms) {
            H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_terms), H->owned.end());
            (void)hipFree(H->run_terms);
            H->run_terms = nullptr;
        }
        std::vector<double> tv = H->run_terms_host;
        if (tv.empty()) tv.assign(3, 0.0);
        if ((s = upload(H, &H->run_terms, tv.data(), tv.size()))) return s;
        if (H->run_pix_terms) {
            H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), (void*)H->run_pix_terms), H->owned.end());
            (void)hipFree(H->run_pix_terms);
            H->run_pix_terms = nullptr;
        }
        if (H->src_pix) {
            if (H->run_pix_host.size() != 2 * (size_t)H->npix * (size_t)nscales)
                return fail(1, "run_setup: gridded rain needs w2d_run_pixel_terms for every window first");
            std::vector<double> pv = H->run_pix_host;
            if (pv.empty()) pv.assign(2, 0.0);
            if ((s = upload(H, &H->run_pix_terms, pv.data(), pv.size()))) return s;
        }
    }
    if (num->maxima < 0) return fail(1, "run_setup: maxima is an interval in steps (0: none)");
    H->run_maxima = num->maxima;
    if (H->run_maxima && !H->maxima[0]) return fail(1, "run_setup: maxima asked for but maxima_set not called");
    if (H->graph) {
        (void)hipGraphExecDestroy(H->graph);
        H->graph = nullptr;
        H->graph_steps = 0;
    }
    H->run_m = M;
    H->run_dirty = 0;
    H->members_stale = 0;
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_run_steps(w2d_handle* H, int32_t nsteps, int32_t graph_steps) {
    if (!H || !H->run_m || nsteps < 0 || graph_steps < 0) return fail(1, "run_steps: call run_setup first");
    int done = 0, s;
    if ((s = members_follow(H))) return s;
    // a momentum stored from the host is read by one step launched on its own: captured steps take it as recomputed
    if (H->m_fresh && nsteps > 0) {
        if ((s = run_step_any(H))) return s;
        ++done;
    }
    // HIP cannot time events recorded inside a captured graph: timed steps are launched one by one. Every kernel of a
    // step takes its buffers from its member's parities on the device, so one graph replays at any parity.
    if (graph_steps > 0 && !H->timing) {
        if (H->graph && H->graph_steps != graph_steps) {
            (void)hipGraphExecDestroy(H->graph);
            H->graph = nullptr;
        }
        if (!H->graph && nsteps - done >= graph_steps) {
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(H->stream, hipStreamCaptureModeThreadLocal));
            for (int q = 0; q < graph_steps; ++q)
                if ((s = run_step_any(H))) {
                    (void)hipStreamEndCapture(H->stream, &g);
                    return s;
                }
            HIP_TRY(hipStreamEndCapture(H->stream, &g));
            HIP_TRY(hipGraphInstantiate(&H->graph, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            H->graph_steps = graph_steps;
        }
        while (H->graph && nsteps - done >= graph_steps) {
            HIP_TRY(hipGraphLaunch(H->graph, H->stream));
            done += graph_steps;
        }
        H->run_dirty = 1;
    }
    for (; done < nsteps; ++done)
        if ((s = run_step_any(H))) return s;
    // no update waits past the call: whatever reads the state next reads it whole
    return nsteps > 0 ? run_drain(H) : 0;
}

int w2d_run_pipeline(w2d_handle* H, int32_t on) {
    if (!H || (on != 0 && on != 1)) return fail(1, "run_pipeline: bad arguments");
    H->pipe_on = on;
    if (!H->run_m || (int)H->pipe_ok.size() != H->run_m) return 0;
    // a run set up already takes it from its next step (run_steps returned with no update waiting)
    std::vector<Member> mem;
    int s;
    if ((s = members_read(H, mem))) return s;
    int any = 0;
    for (int m = 0; m < H->run_m; ++m) {
        mem[m].pipe = on && H->pipe_ok[m];
        any |= mem[m].pipe;
    }
    if ((s = members_write(H, mem))) return s;
    if (any != H->run_pipe && H->graph) {  // a captured graph holds the other step's launches
        (void)hipGraphExecDestroy(H->graph);
        H->graph = nullptr;
        H->graph_steps = 0;
    }
    H->run_pipe = any;
    return 0;
}

int w2d_run_pause(w2d_handle* H, const double* t_stop, int32_t M) {
    if (!H || !t_stop || M != H->run_m) return fail(1, "run_pause: bad arguments");
    std::vector<Member> mem(M);
    HIP_TRY(hipMemcpyAsync(mem.data(), H->run_mem, M * sizeof(Member), hipMemcpyDeviceToHost, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    for (int m = 0; m < M; ++m) mem[m].t_stop = t_stop[m];
    HIP_TRY(hipMemcpyAsync(H->run_mem, mem.data(), M * sizeof(Member), hipMemcpyHostToDevice, H->stream));
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_run_read(w2d_handle* H, w2d_member_state* out, int32_t M) {
    if (!H || !out || M != H->run_m) return fail(1, "run_read: bad arguments");
    H->run_dirty = 1;  // a check the last step asked for is applied now, as the host loop does after its step
    int s;
    if ((s = run_align(H))) return s;
    std::vector<Member> mem;
    if ((s = members_read(H, mem))) return s;
    for (int m = 0; m < M; ++m) {
        const Member& b = mem[m];
        w2d_member_state& o = out[m];
        o.t = b.t;
        o.dt = b.dt;
        o.nsteps = b.nsteps;
        o.done = b.done;
        o.failed = b.failed;
        o.refresh_due = b.refresh_due;
        o.dry = b.dry;
        o.dt_min = b.dt_min;
        o.q_peak = b.q_peak;
        o.t_peak = b.t_peak;
        o.pipelined = b.pipe;
        o.pipelined_updates = b.u_made;
        // rain credited pro rata in the current window (monitor.py:197-212)
        const double len = b.window_end - b.window_start;
        const double share = len > 0.0 ? (std::min(b.t, b.window_end) - b.window_start) / len : 0.0;
        o.rain = b.rain_done + b.rain_window * share;
    }
    return 0;
}

int w2d_tally_totals(w2d_handle* H, double* out) {
    if (!H || !out) return fail(1, "tally_totals: bad arguments");
    if (H->n_tally) {
        HIP_TRY(hipMemcpyAsync(out, H->t_total, H->n_tally * sizeof(double), hipMemcpyDeviceToHost, H->stream));
        HIP_TRY(hipStreamSynchronize(H->stream));
    }
    return 0;
}

int w2d_active_tiles(w2d_handle* H, int32_t* out, int32_t* n) {
    if (!H || !n) return fail(1, "active_tiles: bad arguments");
    *n = H->nact;
    if (out && H->nact) std::copy(H->active_host.begin(), H->active_host.end(), out);
    return 0;
}

int w2d_run_idle(w2d_handle* H, int32_t* out) {
    if (!H || !out || !H->run_idle) return fail(1, "run_idle: bad arguments");
    HIP_TRY(hipMemcpy(out, H->run_idle, H->ntiles * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

#include "levels_host.hpp"
