// Device side of the FDTD step: mesh view, stress/displacement stencils, the
// naive sweep, the boundary-face kernel, probes, energy audit, pack/unpack.
//
// Layout in HBM (one mesh): each displacement component is a dense
// (nx, ny, pz) array, z contiguous, row stride pz (= nz: unpadded), for two
// time levels (ping-pong); a level's arrays lie back to back in one
// allocation.  Cells outside a component's reference extent (ux has nx-1
// planes, uy ny-1 rows, uz nz-1 columns) are allocated, never written and
// stay zero.  Material: one uint8 index per node in the same layout, and a
// 16-value row per material.
//
// Semantics of one step (what the reference does in base_solver.py:245-256):
//   1. the wall line uz[0, :, 0] takes the source value w(tt); here the value
//      is substituted on every read of that line instead of being stored;
//   2. stresses from u (T is never stored: each stencil recomputes it);
//   3. "stage I": the interior and free-surface displacement update;
//   4. the four Mur faces, x=max, y=0, y=max, z=max, each reading the plane
//      one cell inside as updated so far and overwriting shared edges;
//   5. rotation: the new level becomes current (pointer swap).
// Stage 4 is resolved per cell: the last face that writes a cell fixes its
// value as a chain  face(... face(stageI(inner cell)))  of at most four maps,
// and every stage-I value a chain reads lies strictly inside all faces, where
// the sweep kernel has already written it.  So the sweep writes all non-face
// cells plus the z = max face cells (a single map of the cell below), and a
// second, small kernel writes the other face cells.  Cells of uy/uz on the
// x = 0 wall that the reference never updates are written as 0, which is
// what the reference holds there (the source line included, whose stored
// value after a step is the never-written 0 of u_new).
#pragma once
#include "pf_formulas.hpp"

namespace pf {

template <class R>
struct Mesh {
    int nx, ny, nz, pz;
    long long plane;  // ny * pz
    // level n (current), hi part
    const R* __restrict__ c0;
    const R* __restrict__ c1;
    const R* __restrict__ c2;
    // level written by the step (level n+1), hi part
    R* n0;
    R* n1;
    R* n2;
    // level n-1 (read at the node only; the same arrays as n* when the step
    // overwrites the previous level in place)
    const R* o0;
    const R* o1;
    const R* o2;
    // lo parts (PF_FP32C only), 16-bit (lo_encode / lo_decode)
    const lo_t* __restrict__ cl0;
    const lo_t* __restrict__ cl1;
    const lo_t* __restrict__ cl2;
    lo_t* nl0;
    lo_t* nl1;
    lo_t* nl2;
    const lo_t* ol0;
    const lo_t* ol1;
    const lo_t* ol2;
    const uint8_t* __restrict__ ids;
    const R* __restrict__ mat;
    // spacing tables: FP64 = spacings, FP32 = inverse spacings
    const R* __restrict__ fx;
    const R* __restrict__ fy;
    const R* __restrict__ fz;
    const R* __restrict__ sx;
    const R* __restrict__ sy;
    const R* __restrict__ sz;
    R fx0, fy0, fz0, sx0, sy0, sz0;
    const R* __restrict__ dpx;  // sponge terms per axis (null: none)
    const R* __restrict__ dpy;
    const R* __restrict__ dpz;
    R mur[12];
    R mur_lo[12];  // FP32C: low parts of the Mur coefficients
    const double* __restrict__ wave;
    const long long* __restrict__ step_base;  // device step counter or null
    long long step_off;                       // tt = *step_base + step_off
    long long nsteps;                         // step budget (batch members)
    int nmat;                                 // rows in mat
    int surf;  // 1: the surface u_z update divides T3's difference by fdz (extension); 0: the reference's precedence
};

template <class R>
PF_DEV long long at(const Mesh<R>& m, int i, int j, int k) {
    return (long long)i * m.plane + (long long)j * m.pz + k;
}

template <class R>
PF_DEV long long step_index(const Mesh<R>& m) {
    return (m.step_base ? *m.step_base : 0) + m.step_off;
}

// Component extents: last valid index per axis.
template <int Q> struct Ext;
template <> struct Ext<0> { static PF_DEV int Lx(int nx) { return nx - 2; } static PF_DEV int Ly(int ny) { return ny - 1; } static PF_DEV int Lz(int nz) { return nz - 1; } };
template <> struct Ext<1> { static PF_DEV int Lx(int nx) { return nx - 1; } static PF_DEV int Ly(int ny) { return ny - 2; } static PF_DEV int Lz(int nz) { return nz - 1; } };
template <> struct Ext<2> { static PF_DEV int Lx(int nx) { return nx - 1; } static PF_DEV int Ly(int ny) { return ny - 1; } static PF_DEV int Lz(int nz) { return nz - 2; } };

// ---------------------------------------------------------------- loads ----
template <class R> PF_DEV R lux(const Mesh<R>& m, int i, int j, int k) { return m.c0[at(m, i, j, k)]; }
template <class R> PF_DEV R luy(const Mesh<R>& m, int i, int j, int k) { return m.c1[at(m, i, j, k)]; }
template <class R> PF_DEV R luz(const Mesh<R>& m, int i, int j, int k, R w) {
    R v = m.c2[at(m, i, j, k)];
    return (i == 0 && k == 0) ? w : v;
}
template <class R> PF_DEV const R* matp(const Mesh<R>& m, int i, int j, int k) {
    return m.mat + MAT_STRIDE * (int)m.ids[at(m, i, j, k)];
}

// ------------------------------------------------------------- stresses ----
// Each returns the value the reference's T array holds at (i,j,k) after
// update_T + apply_T_tfbc, and 0 where the reference never writes it.
template <int EXACT, class R>
PF_DEV void tn(const Mesh<R>& m, int i, int j, int k, R w, R& t1, R& t2, R& t3) {
    t1 = t2 = t3 = R(0);
    if (i < 1 || i > m.nx - 2 || j < 1 || j > m.ny - 2 || k < 0 || k > m.nz - 2) return;
    const R* c = matp(m, i, j, k);
    R dux = lux(m, i, j, k) - lux(m, i - 1, j, k);
    R duy = luy(m, i, j, k) - luy(m, i, j - 1, k);
    if (k >= 1) {
        R duz = luz(m, i, j, k, w) - luz(m, i, j, k - 1, w);
        normal3<EXACT>(c, dux, duy, duz, m.sx[i - 1], m.sy[j - 1], m.sz[k - 1], t1, t2, t3);
    } else {
        normal3_surface<EXACT>(c, dux, duy, luz(m, i, j, 0, w), m.sx0, m.sy0, m.sz0, t1, t2, t3);
    }
}

template <int EXACT, class R>
PF_DEV R t4(const Mesh<R>& m, int i, int j, int k, R w) {
    if (i < 1 || i > m.nx - 2 || j < 0 || j > m.ny - 2 || k < 0 || k > m.nz - 2) return R(0);
    if (k >= 1) {
        R c = matp(m, i, j + 1, k + 1)[M_C33];
        return shear<EXACT>(c, luy(m, i, j, k + 1) - luy(m, i, j, k), m.fz[k],
                            luz(m, i, j + 1, k, w) - luz(m, i, j, k, w), m.fy[j]);
    }
    R c = matp(m, i, j + 1, 0)[M_C33];
    return shear<EXACT>(c, luy(m, i, j, 1) - luy(m, i, j, 0), m.fy0,
                        luz(m, i, j + 1, 0, w) - luz(m, i, j, 0, w), m.fz0);
}

template <int EXACT, class R>
PF_DEV R t5(const Mesh<R>& m, int i, int j, int k, R w) {
    if (i < 0 || i > m.nx - 2 || j < 1 || j > m.ny - 2 || k < 0 || k > m.nz - 2) return R(0);
    if (k >= 1) {
        R c = matp(m, i + 1, j, k + 1)[M_C44];
        return shear<EXACT>(c, lux(m, i, j, k + 1) - lux(m, i, j, k), m.fz[k],
                            luz(m, i + 1, j, k, w) - luz(m, i, j, k, w), m.fx[i]);
    }
    R c = matp(m, i + 1, j, 0)[M_C44];
    return shear<EXACT>(c, lux(m, i, j, 1) - lux(m, i, j, 0), m.fx0,
                        luz(m, i + 1, j, 0, w) - luz(m, i, j, 0, w), m.fz0);
}

template <int EXACT, class R>
PF_DEV R t6(const Mesh<R>& m, int i, int j, int k, R w) {
    if (i < 0 || i > m.nx - 2 || j < 0 || j > m.ny - 2 || k < 0 || k > m.nz - 2) return R(0);
    (void)w;
    if (k >= 1) {
        R c = matp(m, i + 1, j + 1, k)[M_C55];
        return shear<EXACT>(c, lux(m, i, j + 1, k) - lux(m, i, j, k), m.fy[j],
                            luy(m, i + 1, j, k) - luy(m, i, j, k), m.fx[i]);
    }
    R c = matp(m, i + 1, j + 1, 0)[M_C55];
    return shear<EXACT>(c, lux(m, i, j + 1, 0) - lux(m, i, j, 0), m.fx0,
                        luy(m, i + 1, j, 0) - luy(m, i, j, 0), m.fz0);
}

template <class R>
PF_DEV R damp_at(const Mesh<R>& m, int i, int j, int k) {
    return m.dpx[i] + m.dpy[j] + m.dpz[k];
}

// Stage-I bracket ("div") and coefficient of each component at a non-face
// cell (base_solver.py:441-463 for k >= 1, :495-517 for k = 0).
template <int EXACT, class R>
PF_DEV void stage_ux(const Mesh<R>& m, int i, int j, int k, R w, R& coef, R& dv) {
    R a1, a2, a3, b1, b2, b3;
    tn<EXACT>(m, i + 1, j, k, w, a1, a2, a3);
    tn<EXACT>(m, i, j, k, w, b1, b2, b3);
    R d6 = t6<EXACT>(m, i, j, k, w) - t6<EXACT>(m, i, j - 1, k, w);
    if (k >= 1) {
        R d5 = t5<EXACT>(m, i, j, k, w) - t5<EXACT>(m, i, j, k - 1, w);
        dv = div3<EXACT>(a1 - b1, m.fx[i], d6, m.sy[j - 1], d5, m.sz[k - 1]);
        coef = matp(m, i + 1, j, k)[M_COEF];
    } else {
        R d5 = t5<EXACT>(m, i, j, 0, w) - R(0);
        dv = div3<EXACT>(a1 - b1, m.fx0, d6, m.sy0, d5, m.sz0);
        coef = matp(m, i + 1, j, 0)[M_COEF];
    }
}

template <int EXACT, class R>
PF_DEV void stage_uy(const Mesh<R>& m, int i, int j, int k, R w, R& coef, R& dv) {
    R d6 = t6<EXACT>(m, i, j, k, w) - t6<EXACT>(m, i - 1, j, k, w);
    R a1, a2, a3, b1, b2, b3;
    tn<EXACT>(m, i, j + 1, k, w, a1, a2, a3);
    tn<EXACT>(m, i, j, k, w, b1, b2, b3);
    if (k >= 1) {
        R d4 = t4<EXACT>(m, i, j, k, w) - t4<EXACT>(m, i, j, k - 1, w);
        dv = div3<EXACT>(d6, m.sx[i - 1], a2 - b2, m.fy[j], d4, m.sz[k - 1]);
        coef = matp(m, i, j + 1, k)[M_COEF];
    } else {
        R d4 = t4<EXACT>(m, i, j, 0, w) - R(0);
        dv = div3<EXACT>(d6, m.sx0, a2 - b2, m.fy0, d4, m.sz0);
        coef = matp(m, i, j + 1, 0)[M_COEF];
    }
}

template <int EXACT, class R>
PF_DEV void stage_uz(const Mesh<R>& m, int i, int j, int k, R w, R& coef, R& dv) {
    R d5 = t5<EXACT>(m, i, j, k, w) - t5<EXACT>(m, i - 1, j, k, w);
    R d4 = t4<EXACT>(m, i, j, k, w) - t4<EXACT>(m, i, j - 1, k, w);
    R a1, a2, a3, b1, b2, b3;
    tn<EXACT>(m, i, j, k + 1, w, a1, a2, a3);
    tn<EXACT>(m, i, j, k, w, b1, b2, b3);
    if (k >= 1) {
        dv = div3<EXACT>(d5, m.sx[i - 1], d4, m.sy[j - 1], a3 - b3, m.fz[k]);
        coef = matp(m, i, j, k + 1)[M_COEF];
    } else {
        dv = m.surf ? div3<EXACT>(d5, m.sx0, d4, m.sy0, a3 - b3, m.fz0)
                    : div3_uz_surface<EXACT>(d5, m.sx0, d4, m.sy0, a3, b3, m.fz0);
        coef = matp(m, i, j, 0)[M_COEF];  // reference samples rho at k = 0 here
    }
}

// ----------------------------------------------------------- faces ----
// A value in flight through the face chain: hi (+ lo in compensated mode).
template <class R> struct V { R h, l; };

template <int Q, class R>
PF_DEV V<R> cur_q(const Mesh<R>& m, int i, int j, int k, R w) {
    long long id = at(m, i, j, k);
    const R* c = Q == 0 ? m.c0 : (Q == 1 ? m.c1 : m.c2);
    const lo_t* cl = Q == 0 ? m.cl0 : (Q == 1 ? m.cl1 : m.cl2);
    V<R> v;
    v.h = c[id];
    v.l = cl ? lo_decode(v.h, cl[id]) : R(0);
    if (Q == 2 && i == 0 && k == 0) { v.h = w; v.l = R(0); }
    return v;
}

template <int Q, class R>
PF_DEV V<R> new_q(const Mesh<R>& m, int i, int j, int k) {
    long long id = at(m, i, j, k);
    const R* n = Q == 0 ? m.n0 : (Q == 1 ? m.n1 : m.n2);
    const lo_t* nl = Q == 0 ? m.nl0 : (Q == 1 ? m.nl1 : m.nl2);
    V<R> v;
    v.h = n[id];
    v.l = nl ? lo_decode(v.h, nl[id]) : R(0);
    return v;
}

// a + c * (b - d), the reference's `q[..] + c*(q_new[..] - q[..])`
template <int MODE, class R>
PF_DEV V<R> mur_map(V<R> a, R ch, R cl, V<R> b, V<R> d) {
    V<R> r;
    if constexpr (MODE == MODE_FP32C) {
        mur_comp(a.h, a.l, ch, cl, b.h, b.l, d.h, d.l, r.h, r.l);
    } else {
        r.h = a.h + ch * (b.h - d.h);
        r.l = R(0);
    }
    return r;
}

// Value after stage I and the x = max face (face index 0).
template <int Q, int MODE, class R>
PF_DEV V<R> after_x(const Mesh<R>& m, int i, int j, int k, R w) {
    const int Lx = Ext<Q>::Lx(m.nx);
    if (i == Lx)
        return mur_map<MODE>(cur_q<Q>(m, Lx - 1, j, k, w), m.mur[0 + Q], m.mur_lo[0 + Q],
                             new_q<Q>(m, Lx - 1, j, k), cur_q<Q>(m, Lx, j, k, w));
    return new_q<Q>(m, i, j, k);
}
// ... then the y = 0 face (face index 1)
template <int Q, int MODE, class R>
PF_DEV V<R> after_y0(const Mesh<R>& m, int i, int j, int k, R w) {
    if (j == 0)
        return mur_map<MODE>(cur_q<Q>(m, i, 1, k, w), m.mur[3 + Q], m.mur_lo[3 + Q],
                             after_x<Q, MODE>(m, i, 1, k, w), cur_q<Q>(m, i, 0, k, w));
    return after_x<Q, MODE>(m, i, j, k, w);
}
// ... then the y = max face (face index 2)
template <int Q, int MODE, class R>
PF_DEV V<R> after_y1(const Mesh<R>& m, int i, int j, int k, R w) {
    const int Ly = Ext<Q>::Ly(m.ny);
    if (j == Ly)
        return mur_map<MODE>(cur_q<Q>(m, i, Ly - 1, k, w), m.mur[6 + Q], m.mur_lo[6 + Q],
                             after_y0<Q, MODE>(m, i, Ly - 1, k, w), cur_q<Q>(m, i, Ly, k, w));
    return after_y0<Q, MODE>(m, i, j, k, w);
}
// ... then the z = max face (face index 3): the final value.
template <int Q, int MODE, class R>
PF_DEV V<R> final_value(const Mesh<R>& m, int i, int j, int k, R w) {
    const int Lz = Ext<Q>::Lz(m.nz);
    if (k == Lz)
        return mur_map<MODE>(cur_q<Q>(m, i, j, Lz - 1, w), m.mur[9 + Q], m.mur_lo[9 + Q],
                             after_y1<Q, MODE>(m, i, j, Lz - 1, w), cur_q<Q>(m, i, j, Lz, w));
    return after_y1<Q, MODE>(m, i, j, k, w);
}

template <int Q, class R> PF_DEV R* nptr(const Mesh<R>& m) { return Q == 0 ? m.n0 : (Q == 1 ? m.n1 : m.n2); }
template <int Q, class R> PF_DEV lo_t* nlptr(const Mesh<R>& m) { return Q == 0 ? m.nl0 : (Q == 1 ? m.nl1 : m.nl2); }

template <int Q, int MODE, class R>
PF_DEV void put(const Mesh<R>& m, int i, int j, int k, V<R> v) {
    const long long id = at(m, i, j, k);
    nptr<Q>(m)[id] = v.h;
    if constexpr (MODE == MODE_FP32C) nlptr<Q>(m)[id] = lo_encode(v.h, v.l);
}

// Write the stage-I value s0 of a non-face cell, and, when k = Lz-1, the
// z = max face cell above it, whose final value is a single Mur map of s0
// (there the data is already in the sweep's registers; a separate pass
// would read it back with a stride of one row per cell).  `cur` is the
// cell's own level-n value.
template <int Q, int MODE, class R>
PF_DEV void put_with_zface(const Mesh<R>& m, int i, int j, int k, V<R> s0, V<R> cur, R w) {
    put<Q, MODE>(m, i, j, k, s0);
    const int Lz = Ext<Q>::Lz(m.nz);
    if (k == Lz - 1)
        put<Q, MODE>(m, i, j, Lz, mur_map<MODE>(cur, m.mur[9 + Q], m.mur_lo[9 + Q], s0, cur_q<Q>(m, i, j, Lz, w)));
}

// Face cells of component Q not written by the sweep, once each, z fastest
// so each run of threads touches contiguous memory:
//   X  : (Lx, j, k)   0 < j < Ly, k < Lz
//   Y0 : (i, 0, k)    i < Lx,     k < Lz
//   Y1 : (i, Ly, k)   i < Lx,     k < Lz
//   box edges (two or more faces meet):
//        (Lx, 0, k), (Lx, Ly, k)   all k
//        (Lx, j, Lz)               0 < j < Ly
//        (i, 0, Lz), (i, Ly, Lz)   i < Lx
template <int Q>
PF_DEV long long face_count(int nx, int ny, int nz) {
    const long long Lx = Ext<Q>::Lx(nx), Ly = Ext<Q>::Ly(ny), Lz = Ext<Q>::Lz(nz);
    return (Ly - 1) * Lz + 2 * Lx * Lz + 2 * (Lz + 1) + (Ly - 1) + 2 * Lx;
}

template <int Q>
PF_DEV void face_cell(int nx, int ny, int nz, long long t, int& i, int& j, int& k) {
    const long long Lx = Ext<Q>::Lx(nx), Ly = Ext<Q>::Ly(ny), Lz = Ext<Q>::Lz(nz);
    long long n = (Ly - 1) * Lz;
    if (t < n) { i = (int)Lx; j = 1 + (int)(t / Lz); k = (int)(t % Lz); return; }
    t -= n;
    n = Lx * Lz;
    if (t < n) { j = 0; i = (int)(t / Lz); k = (int)(t % Lz); return; }
    t -= n;
    if (t < n) { j = (int)Ly; i = (int)(t / Lz); k = (int)(t % Lz); return; }
    t -= n;
    if (t < Lz + 1) { i = (int)Lx; j = 0; k = (int)t; return; }
    t -= Lz + 1;
    if (t < Lz + 1) { i = (int)Lx; j = (int)Ly; k = (int)t; return; }
    t -= Lz + 1;
    if (t < Ly - 1) { i = (int)Lx; j = 1 + (int)t; k = (int)Lz; return; }
    t -= Ly - 1;
    if (t < Lx) { i = (int)t; j = 0; k = (int)Lz; return; }
    t -= Lx;
    i = (int)t; j = (int)Ly; k = (int)Lz;
}

template <int Q, int MODE, class R>
PF_DEV void face_write(const Mesh<R>& m, long long t, R w) {
    int i, j, k;
    face_cell<Q>(m.nx, m.ny, m.nz, t, i, j, k);
    put<Q, MODE>(m, i, j, k, final_value<Q, MODE>(m, i, j, k, w));
}

// Face cell t (all three components' cells numbered one after another) of a
// step whose source value is w.
template <class R, int MODE>
PF_DEV void faces_one(const Mesh<R>& m, long long t, R w) {
    const long long e0 = face_count<0>(m.nx, m.ny, m.nz);
    const long long e1 = face_count<1>(m.nx, m.ny, m.nz);
    const long long e2 = face_count<2>(m.nx, m.ny, m.nz);
    if (t < e0) { face_write<0, MODE>(m, t, w); return; }
    t -= e0;
    if (t < e1) { face_write<1, MODE>(m, t, w); return; }
    t -= e1;
    if (t < e2) face_write<2, MODE>(m, t, w);
}

// Face cells [t0, t0 + 256 * PER) of view m, PER per thread of a
// 256-thread block; the step index and source value are read once.  One
// mesh's faces run best with one cell per thread (more blocks in flight);
// the batch, whose blocks first read their member's tables, with four.
template <class R, int MODE, int PER>
PF_DEV void faces_run(const Mesh<R>& m, long long t0, long long total) {
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;
    const R w = (R)m.wave[tt];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const long long t = t0 + r * 256 + threadIdx.x;
        if (t < total) faces_one<R, MODE>(m, t, w);
    }
}

template <class R, int MODE>
__global__ __launch_bounds__(256) void k_faces(Mesh<R> m, long long total) {
    faces_run<R, MODE, 1>(m, (long long)blockIdx.x * 256, total);
}

// The same face cells, arranged for one mesh's step.  The rows of the three
// face planes
//   X  (Lx, j, k)  0 < j < Ly,   Y0 (i, 0, k), Y1 (i, Ly, k)  i < Lx,   k < Lz
// take one Mur map of interior cells each:
//   X  map(cur(Lx-1, j, k), new(Lx-1, j, k), cur(Lx, j, k))
//   Y0 map(cur(i, 1, k),    new(i, 1, k),    cur(i, 0, k))
//   Y1 map(cur(i, Ly-1, k), new(i, Ly-1, k), cur(i, Ly, k))
// (final_value's chain, with after_x and after_y0 reduced to the stage-I
// value new(), which holds for i < Lx and Ly - 1 > 0).  A wave takes a run
// of FACE_RUN cells of one row, FACE_LANE_CELLS a lane, and issues all its
// loads before it stores anything, so one round of memory latency covers
// them (eight a lane measured no faster than four: the launch is bound by
// its fixed cost, about 4 us of a 9.6 us launch at 513 x 513 x 257).  The
// box edges (the rest of face_cell's list, a few thousand cells) go one a
// thread through final_value, in the last blocks (measured 0.4 us faster
// per launch than putting them first).  Needs Ly >= 2 for every
// component.
constexpr int FACE_LANE_CELLS = 4;                 // cells of a row per lane
constexpr int FACE_RUN = 64 * FACE_LANE_CELLS;     // cells of a row per wave
PF_DEV int face_runs(int nz) { return (nz - 1 + FACE_RUN - 1) / FACE_RUN; }

template <int Q>
PF_DEV int face_rows(int nx, int ny) {
    return Ext<Q>::Ly(ny) - 1 + 2 * Ext<Q>::Lx(nx);
}

template <int Q, int MODE, class R>
PF_DEV void face_row_run(const Mesh<R>& m, int r, int run, int lane, R w) {
    const int Lx = Ext<Q>::Lx(m.nx), Ly = Ext<Q>::Ly(m.ny), Lz = Ext<Q>::Lz(m.nz);
    int i, j, ii, ji;  // the face row, the interior row its map reads
    R mh, ml;          // its Mur coefficient (constant indices: a run-time index
                       // into the kernel argument would copy it to scratch)
    if (r < Ly - 1) {
        i = Lx; j = 1 + r; ii = Lx - 1; ji = j; mh = m.mur[0 + Q]; ml = m.mur_lo[0 + Q];
    } else if (r < Ly - 1 + Lx) {
        i = r - (Ly - 1); j = 0; ii = i; ji = 1; mh = m.mur[3 + Q]; ml = m.mur_lo[3 + Q];
    } else {
        i = r - (Ly - 1) - Lx; j = Ly; ii = i; ji = Ly - 1; mh = m.mur[6 + Q]; ml = m.mur_lo[6 + Q];
    }
    const R* c = Q == 0 ? m.c0 : (Q == 1 ? m.c1 : m.c2);
    const R* n = Q == 0 ? m.n0 : (Q == 1 ? m.n1 : m.n2);
    const lo_t* cl = Q == 0 ? m.cl0 : (Q == 1 ? m.cl1 : m.cl2);
    const lo_t* nl = Q == 0 ? m.nl0 : (Q == 1 ? m.nl1 : m.nl2);
    const long long fb = at(m, i, j, 0), ib = at(m, ii, ji, 0);
    V<R> a[FACE_LANE_CELLS], b[FACE_LANE_CELLS], d[FACE_LANE_CELLS];
#pragma unroll
    for (int u = 0; u < FACE_LANE_CELLS; u++) {
        const int k = run * FACE_RUN + u * 64 + lane;
        const int kk = k < Lz ? k : Lz - 1;  // lanes past the row load its last cell and store nothing
        a[u].h = c[ib + kk];
        b[u].h = n[ib + kk];
        d[u].h = c[fb + kk];
        if constexpr (MODE == MODE_FP32C) {
            a[u].l = lo_decode(a[u].h, cl[ib + kk]);
            b[u].l = lo_decode(b[u].h, nl[ib + kk]);
            d[u].l = lo_decode(d[u].h, cl[fb + kk]);
        } else {
            a[u].l = b[u].l = d[u].l = R(0);
        }
    }
#pragma unroll
    for (int u = 0; u < FACE_LANE_CELLS; u++) {
        const int k = run * FACE_RUN + u * 64 + lane;
        if (k >= Lz) continue;
        if (Q == 2 && k == 0) {  // the source line uz[0, :, 0] reads as w
            if (ii == 0) { a[u].h = w; a[u].l = R(0); }
            if (i == 0) { d[u].h = w; d[u].l = R(0); }
        }
        put<Q, MODE>(m, i, j, k, mur_map<MODE>(a[u], mh, ml, b[u], d[u]));
    }
}

template <int Q>
PF_DEV long long face_edge_base(int nx, int ny, int nz) {
    const long long Lx = Ext<Q>::Lx(nx), Ly = Ext<Q>::Ly(ny), Lz = Ext<Q>::Lz(nz);
    return (Ly - 1) * Lz + 2 * Lx * Lz;
}

// A probe frame recorded by the face launch (probe blocks after the row
// and edge blocks; blocks = 0: none).  The frame index advances as in
// k_probe.
struct FaceProbe {
    double* ring;
    long long* frame;
    int* done;
    int i0, si, ni, j0, sj, nj, cap, every, blocks;
};

// Probe point p of the u_z surface (k = 0) of the level m writes, as a later
// read would see it.  Points on a u_z face (i = Lx, j = 0 or Ly) are not
// read: the face blocks of the same launch store them, so the point takes
// the same chain value (final_value) with its lo part rounded as stored.
template <class R, int MODE>
PF_DEV double probe_point(const Mesh<R>& m, const FaceProbe& pr, int p, R w) {
    const int i = pr.i0 + pr.si * (p / pr.nj), j = pr.j0 + pr.sj * (p % pr.nj);
    if (i == Ext<2>::Lx(m.nx) || j == 0 || j == Ext<2>::Ly(m.ny)) {
        const V<R> v = final_value<2, MODE>(m, i, j, 0, w);
        double d = (double)v.h;
        if constexpr (MODE == MODE_FP32C) d += (double)lo_decode(v.h, lo_encode(v.h, v.l));
        return d;
    }
    const long long id = at(m, i, j, 0);
    const R h = m.n2[id];
    double d = (double)h;
    if (m.nl2) d += (double)lo_decode(h, m.nl2[id]);
    return d;
}

// Row block rb of k_faces_rows' grid for view m (four runs, a wave each).
template <class R, int MODE>
PF_DEV void face_rows_block(const Mesh<R>& m, int rb, R w) {
    const int kr = face_runs(m.nz);
    const int lane = threadIdx.x % 64;
    int t = rb * 4 + (int)threadIdx.x / 64;  // this wave's run
    const int w0 = face_rows<0>(m.nx, m.ny) * kr;
    if (t < w0) { face_row_run<0, MODE>(m, t / kr, t % kr, lane, w); return; }
    t -= w0;
    const int w1 = face_rows<1>(m.nx, m.ny) * kr;
    if (t < w1) { face_row_run<1, MODE>(m, t / kr, t % kr, lane, w); return; }
    t -= w1;
    if (t < face_rows<2>(m.nx, m.ny) * kr) face_row_run<2, MODE>(m, t / kr, t % kr, lane, w);
}

// Edge block eb (256 box-edge cells, one a thread) for view m.
template <class R, int MODE>
PF_DEV void face_edges_block(const Mesh<R>& m, int eb, R w) {
    long long e = (long long)eb * 256 + threadIdx.x;
    const long long e0 = face_count<0>(m.nx, m.ny, m.nz) - face_edge_base<0>(m.nx, m.ny, m.nz);
    if (e < e0) { face_write<0, MODE>(m, face_edge_base<0>(m.nx, m.ny, m.nz) + e, w); return; }
    e -= e0;
    const long long e1 = face_count<1>(m.nx, m.ny, m.nz) - face_edge_base<1>(m.nx, m.ny, m.nz);
    if (e < e1) { face_write<1, MODE>(m, face_edge_base<1>(m.nx, m.ny, m.nz) + e, w); return; }
    e -= e1;
    const long long e2 = face_count<2>(m.nx, m.ny, m.nz) - face_edge_base<2>(m.nx, m.ny, m.nz);
    if (e < e2) face_write<2, MODE>(m, face_edge_base<2>(m.nx, m.ny, m.nz) + e, w);
}

template <class R, int MODE>
__global__ __launch_bounds__(256) void k_faces_rows(Mesh<R> m, int row_blocks, int edge_blocks, FaceProbe pr) {
    __shared__ int s_last;
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;
    const R w = (R)m.wave[tt];
    if ((int)blockIdx.x >= row_blocks + edge_blocks) {
        // probe blocks: the frame of a recording step (the same for every block)
        if ((tt + 1) % pr.every != 0) return;
        const long long f = *pr.frame;
        const int p = ((int)blockIdx.x - row_blocks - edge_blocks) * 256 + (int)threadIdx.x;
        if (p < pr.ni * pr.nj) pr.ring[(f % pr.cap) * (long long)pr.ni * pr.nj + p] = probe_point<R, MODE>(m, pr, p, w);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            s_last = atomicAdd(pr.done, 1) == pr.blocks - 1;
        }
        __syncthreads();
        if (s_last && threadIdx.x == 0) {
            *pr.frame = f + 1;
            *pr.done = 0;
        }
        return;
    }
    if ((int)blockIdx.x < row_blocks) face_rows_block<R, MODE>(m, (int)blockIdx.x, w);
    else face_edges_block<R, MODE>(m, (int)blockIdx.x - row_blocks, w);
}

// Blocks of k_faces_rows for an nx x ny x nz mesh: row_blocks (four runs
// each), then edge_blocks; false when the mesh is too small for it.
inline bool face_rows_grid(int nx, int ny, int nz, long long& row_blocks, long long& edge_blocks) {
    if (nx < 3 || ny < 4 || nz < 3) return false;
    const long long kr = (nz - 1 + FACE_RUN - 1) / FACE_RUN;
    long long runs = 0, edges = 0;
    const long long L[3][3] = {{nx - 2, ny - 1, nz - 1}, {nx - 1, ny - 2, nz - 1}, {nx - 1, ny - 1, nz - 2}};
    for (const auto& e : L) {
        runs += (e[1] - 1 + 2 * e[0]) * kr;
        edges += 2 * (e[2] + 1) + (e[1] - 1) + 2 * e[0];
    }
    row_blocks = (runs + 3) / 4;
    edge_blocks = (edges + 255) / 256;
    return true;
}

inline long long face_total(int nx, int ny, int nz) {
    auto cnt = [](long long Lx, long long Ly, long long Lz) {
        return (Ly - 1) * Lz + 2 * Lx * Lz + 2 * (Lz + 1) + (Ly - 1) + 2 * Lx;
    };
    return cnt(nx - 2, ny - 1, nz - 1) + cnt(nx - 1, ny - 2, nz - 1) + cnt(nx - 1, ny - 1, nz - 2);
}

// Stage-I value of one component at a cell: the leapfrog of its current
// value cur (hi + lo), its previous value (oh + ol) and coef*div.
template <int MODE, bool SPONGE, class R>
PF_DEV V<R> leap_from(V<R> cur, R oh, R ol, R coef, R dv, R s) {
    V<R> r;
    if constexpr (MODE == MODE_FP32C) {
        leap_comp<SPONGE>(cur.h, cur.l, oh, ol, coef * dv, s, r.h, r.l);
    } else {
        r.h = leap<exact_of(MODE), SPONGE>(cur.h, oh, coef, dv, s);
        r.l = R(0);
    }
    return r;
}

template <int MODE, bool SPONGE, class R>
PF_DEV V<R> leap_value(V<R> cur, const R* nh, const lo_t* nl, long long id, R coef, R dv, R s) {
    const R oh = nh[id];
    return leap_from<MODE, SPONGE>(cur, oh, nl ? lo_decode(oh, nl[id]) : R(0), coef, dv, s);
}

// ------------------------------------------------------- naive sweep ----
// One node (i, j, k) of a step: every stress a stencil needs is recomputed
// from u.  Writes the node's non-face cells and the z = max face cells
// above them.
template <class R, int MODE, bool SPONGE>
PF_DEV void naive_cell(const Mesh<R>& m, int i, int j, int k, R w) {
    constexpr int EXACT = exact_of(MODE);
    const long long id = at(m, i, j, k);
    const R s = SPONGE ? damp_at(m, i, j, k) : R(0);
    // ux: non-face cells i <= nx-3, 1 <= j <= ny-2, k <= nz-2
    if (i <= m.nx - 3 && j >= 1 && j <= m.ny - 2 && k <= m.nz - 2) {
        R coef, dv;
        stage_ux<EXACT>(m, i, j, k, w, coef, dv);
        V<R> cur = cur_q<0>(m, i, j, k, w);
        put_with_zface<0, MODE>(m, i, j, k, leap_value<MODE, SPONGE>(cur, m.o0, m.ol0, id, coef, dv, s), cur, w);
    }
    // uy: non-face cells i <= nx-2, 1 <= j <= ny-3, k <= nz-2; wall i = 0 -> 0
    if (i <= m.nx - 2 && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2) {
        V<R> cur = cur_q<1>(m, i, j, k, w);
        V<R> s0 = {R(0), R(0)};
        if (i > 0) {
            R coef, dv;
            stage_uy<EXACT>(m, i, j, k, w, coef, dv);
            s0 = leap_value<MODE, SPONGE>(cur, m.o1, m.ol1, id, coef, dv, s);
        }
        put_with_zface<1, MODE>(m, i, j, k, s0, cur, w);
    }
    // uz: non-face cells i <= nx-2, 1 <= j <= ny-2, k <= nz-3; wall i = 0 -> 0
    if (i <= m.nx - 2 && j >= 1 && j <= m.ny - 2 && k <= m.nz - 3) {
        V<R> cur = cur_q<2>(m, i, j, k, w);
        V<R> s0 = {R(0), R(0)};
        if (i > 0) {
            R coef, dv;
            stage_uz<EXACT>(m, i, j, k, w, coef, dv);
            s0 = leap_value<MODE, SPONGE>(cur, m.o2, m.ol2, id, coef, dv, s);
        }
        put_with_zface<2, MODE>(m, i, j, k, s0, cur, w);
    }
}

// One thread per node.
template <class R, int MODE, bool SPONGE>
__global__ __launch_bounds__(256) void k_sweep_naive(Mesh<R> m) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= m.nz || j >= m.ny || i >= m.nx) return;
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;
    naive_cell<R, MODE, SPONGE>(m, i, j, k, (R)m.wave[tt]);
}

// ------------------------------------------------------------ probes ----
// One probe frame (surface u_z subgrid, hi + lo) into slot `frame % cap` of
// the ring, after step tt, when (tt + 1) is a multiple of `every`.  Called
// by all threads of one block; thread 0 advances the frame counter after
// every read by the block barrier.
template <class R>
PF_DEV void probe_record(const R* __restrict__ uz, const lo_t* __restrict__ uzl, long long plane, int pz,
                         int i0, int si, int ni, int j0, int sj, int nj, double* __restrict__ ring, int cap,
                         long long* __restrict__ frame, long long tt, int every, long long nsteps) {
    if (tt >= nsteps || (tt + 1) % every != 0) return;
    long long f = *frame;
    double* out = ring + (f % cap) * (long long)ni * nj;
    for (int p = threadIdx.x; p < ni * nj; p += blockDim.x) {
        int a = p / nj, b = p % nj;
        long long id = (long long)(i0 + si * a) * plane + (long long)(j0 + sj * b) * pz;
        const R h = uz[id];
        double v = (double)h;
        if (uzl) v += (double)lo_decode(h, uzl[id]);
        out[p] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) *frame = f + 1;
}

// One probe frame of a single mesh, spread over the blocks of the launch:
// block b gathers points [256 b, 256 b + 256).  Every block reads the frame
// index, writes its points and counts itself done; the last block to count
// (after every other block has read the index) advances it and clears the
// count for the next frame.  No block waits for another.
template <class R>
__global__ __launch_bounds__(256) void k_probe(const R* __restrict__ uz, const lo_t* __restrict__ uzl,
                                               long long plane, int pz, int i0, int si, int ni, int j0, int sj,
                                               int nj, double* __restrict__ ring, int cap,
                                               long long* __restrict__ frame, int* __restrict__ done,
                                               const long long* __restrict__ step_base, long long step_off,
                                               int every, long long nsteps) {
    __shared__ int s_last;
    const long long tt = (step_base ? *step_base : 0) + step_off;  // step just completed
    if (tt >= nsteps || (tt + 1) % every != 0) return;             // the same for every block
    const long long f = *frame;
    double* out = ring + (f % cap) * (long long)ni * nj;
    const int p = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (p < ni * nj) {
        const int a = p / nj, b = p % nj;
        const long long id = (long long)(i0 + si * a) * plane + (long long)(j0 + sj * b) * pz;
        const R h = uz[id];
        double v = (double)h;
        if (uzl) v += (double)lo_decode(h, uzl[id]);
        out[p] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        *frame = f + 1;
        *done = 0;
    }
}

// Tiny kernel: advance a device step counter after a captured chunk.
static __global__ void k_advance(long long* counter, long long by, long long cap) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        long long v = *counter + by;
        *counter = v < cap ? v : cap;
    }
}

// ------------------------------------------------------- pack/unpack ----
// Dense reference-shaped double array <-> padded device component.
template <class R>
__global__ void k_pack(const R* __restrict__ hi, const lo_t* __restrict__ lo, long long plane, int pz,
                       int ex, int ey, int ez, double* __restrict__ out,
                       const double* __restrict__ wsub) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = (long long)ex * ey * ez;
    if (t >= n) return;
    int k = (int)(t % ez);
    long long r = t / ez;
    int j = (int)(r % ey);
    int i = (int)(r / ey);
    long long id = (long long)i * plane + (long long)j * pz + k;
    const R h = hi[id];
    double v = (double)h;
    if (lo) v += (double)lo_decode(h, lo[id]);
    if (wsub && i == 0 && k == 0) v = *wsub;  // the source value on the wall line
    out[t] = v;
}

template <class R>
__global__ void k_unpack(R* __restrict__ hi, lo_t* __restrict__ lo, long long plane, int pz,
                         int ex, int ey, int ez, const double* __restrict__ in) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = (long long)ex * ey * ez;
    if (t >= n) return;
    int k = (int)(t % ez);
    long long r = t / ez;
    int j = (int)(r % ey);
    int i = (int)(r / ey);
    long long id = (long long)i * plane + (long long)j * pz + k;
    double v = in[t];
    R h = (R)v;
    hi[id] = h;
    if (lo) lo[id] = lo_encode(h, (R)(v - (double)h));
}

}  // namespace pf
