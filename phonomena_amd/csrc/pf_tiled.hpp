// LDS-tiled sweep: (y, z) tiles, register march along x.
//
// A block is 64 lanes along z (one wavefront per row) by TJ rows along y and
// marches through a chunk of x planes.  Tiles overlap by one column / row on
// each side: lanes 0 and 63 and rows 0 and TJ-1 compute the stresses their
// neighbours need but write nothing, so every thread runs the same code and
// no halo special-casing is needed.  Per plane each thread
//   1. takes its own column of u at plane i+1 (loaded two planes earlier)
//      and its material index into registers and LDS;
//   2. computes the normal stresses N = (T1, T2, T3) at plane i+1 and the
//      shear stresses T4, T5, T6 at plane i from register values, LDS
//      neighbours along y and DPP lane shifts along z;
//   3. publishes the stresses its y neighbours need through LDS;
//   4. forms the displacement update of plane i (stage I) and writes it,
//      plus the z = max Mur face cell above it when it is the top cell.
// Own-column values of the previous plane (u, N, T5, T6) stay in registers,
// so every u value is read from HBM once per step and written once.
//
// Latency: the march is unrolled by two planes with one register set per
// parity, so the loads of plane i+3 (current level) and of plane i+2
// (previous level) are issued while plane i is computed and no register
// holding a load in flight is ever copied.  Nothing inside the loop reads
// global memory with a per-plane wait: the per-plane x spacings (and sponge
// terms) are staged in LDS once per block.
//
// LDS hazards: su/sid of a parity are rewritten at the top of the plane two
// planes later, before that plane's first barrier, so they are only read
// between a plane's first and second barrier; sn2/ss4/ss6 are rewritten
// after a first barrier and may be read after the second.
//
// The arithmetic is the shared per-value formulas of pf_formulas.hpp, fed
// the same operands as the naive kernel, so both sweeps produce identical
// bits.  A chunk starting at plane ib first runs one warm-up plane (ib-1)
// that only builds the register/LDS state.
#pragma once
#include <type_traits>

#include "pf_runtime.hpp"

namespace pf {

constexpr int TK = 64;        // lanes along z
constexpr int KOUT = TK - 2;  // output columns per tile
constexpr int XT = 256;       // x-table entries: planes per chunk + 3

struct TileMap {
    int nkt, njt, nxc, xchunk;
};

// Block index -> tile index, giving each of the 8 XCDs (which take blocks
// round-robin) one contiguous run of tiles, so tiles that share halo rows
// run on the same L2.
PF_DEV int xcd_remap(int b, int nb) {
    const int q = nb / 8, rem = nb % 8;
    const int x = b % 8, l = b / 8;
    return x * q + (x < rem ? x : rem) + l;
}

// DPP whole-wavefront shifts (gfx9 family: wave_shr:1 / wave_shl:1).  A lane
// with no source keeps its own value, as __shfl_up/__shfl_down do; only
// halo lanes see that.
PF_DEV int dpp_from_below(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
PF_DEV int dpp_from_above(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
PF_DEV float lane_below(float v) { return __int_as_float(dpp_from_below(__float_as_int(v))); }
PF_DEV float lane_above(float v) { return __int_as_float(dpp_from_above(__float_as_int(v))); }
PF_DEV double lane_below(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_below((int)(b & 0xffffffffLL)), hi = dpp_from_below((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
PF_DEV double lane_above(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_above((int)(b & 0xffffffffLL)), hi = dpp_from_above((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Plane accesses go through buffer resources: one SGPR descriptor per array
// (built once per block), the plane selected by the scalar byte offset
// ip * plane * sizeof(R) and the cell by a per-lane 32-bit byte offset that
// does not change along the march, so a load or store costs no per-lane
// address arithmetic.  Arrays must be smaller than 4 GiB (the launcher
// checks); every access made here lies inside its array.
typedef __amdgpu_buffer_rsrc_t Rsrc;
// descriptor word 3 for untyped (raw) buffer access on gfx9-family GPUs
constexpr int PF_RSRC_FLAGS = 0x00020000;

template <class T> PF_DEV Rsrc arr_rsrc(const T* base, long long elems) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, (int)(unsigned)(elems * (long long)sizeof(T)),
                                             PF_RSRC_FLAGS);
}
PF_DEV float bld(Rsrc r, unsigned off, unsigned soff, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
}
PF_DEV double bld(Rsrc r, unsigned off, unsigned soff, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
PF_DEV int bld_u8(Rsrc r, unsigned off, unsigned soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b8(r, (int)off, (int)soff, 0);
}
PF_DEV void bst(Rsrc r, unsigned off, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, (int)soff, 0);
}
PF_DEV void bst(Rsrc r, unsigned off, unsigned soff, double v) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, (int)soff, 0);
}

// A block-uniform value kept in a scalar register (plane offsets).
PF_DEV unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// Descriptors of every array the sweep streams.
struct Bufs {
    Rsrc c0, c1, c2, n0, n1, n2, cl0, cl1, cl2, nl0, nl1, nl2, ids;
};
template <class R> PF_DEV Bufs make_bufs(const Mesh<R>& m) {
    const long long n = (long long)m.nx * m.plane;
    Bufs b;
    b.c0 = arr_rsrc(m.c0, n); b.c1 = arr_rsrc(m.c1, n); b.c2 = arr_rsrc(m.c2, n);
    b.n0 = arr_rsrc(m.n0, n); b.n1 = arr_rsrc(m.n1, n); b.n2 = arr_rsrc(m.n2, n);
    b.cl0 = arr_rsrc(m.cl0, n); b.cl1 = arr_rsrc(m.cl1, n); b.cl2 = arr_rsrc(m.cl2, n);
    b.nl0 = arr_rsrc(m.nl0, n); b.nl1 = arr_rsrc(m.nl1, n); b.nl2 = arr_rsrc(m.nl2, n);
    b.ids = arr_rsrc(m.ids, n);
    return b;
}

// Own column of one plane: u components and material index.
template <class R>
struct Col {
    R x, y, z;
    int id;
};

// The same, as loaded: not yet masked for out-of-mesh lanes nor given the
// source value.  Kept separate so that no select waits on a load in flight.
template <class R>
struct Raw {
    R x, y, z;
    uint8_t id;
};

// Current-level column of plane p, which the caller keeps inside the mesh.
// Loads are never conditional, so that no register with a load in flight is
// merged with another value (which would force a wait and a copy).
template <class R>
PF_DEV Raw<R> load_raw(const Bufs& bf, int p, unsigned plane, unsigned boff, unsigned col) {
    Raw<R> r;
    const unsigned soff = uni((unsigned)p * plane * (unsigned)sizeof(R));
    r.x = bld(bf.c0, boff, soff, R(0));
    r.y = bld(bf.c1, boff, soff, R(0));
    r.z = bld(bf.c2, boff, soff, R(0));
    r.id = __builtin_amdgcn_raw_buffer_load_b8(bf.ids, (int)col, (int)uni((unsigned)p * plane), 0);
    return r;
}

// Out-of-mesh lanes and planes read as 0; the wall line uz(0, j, 0) reads
// the source.
template <class R>
PF_DEV Col<R> settle(const Raw<R>& r, bool valid, bool src, R w) {
    Col<R> c;
    c.x = valid ? r.x : R(0);
    c.y = valid ? r.y : R(0);
    c.z = valid ? (src ? w : r.z) : R(0);
    c.id = valid ? (int)r.id : 0;
    return c;
}

// Values the displacement update of a plane reads at the node itself: the
// previous level (hi, and lo in FP32C) and the current level's lo part.
template <class R>
struct Old {
    R o0, o1, o2, ol0, ol1, ol2, cl0, cl1, cl2;
};

template <class R, bool COMP>
PF_DEV Old<R> load_old(const Bufs& bf, int i, unsigned plane, unsigned boff) {
    const unsigned soff = uni((unsigned)i * plane * (unsigned)sizeof(R));
    Old<R> o;
    o.o0 = bld(bf.n0, boff, soff, R(0));
    o.o1 = bld(bf.n1, boff, soff, R(0));
    o.o2 = bld(bf.n2, boff, soff, R(0));
    if constexpr (COMP) {
        o.ol0 = bld(bf.nl0, boff, soff, R(0));
        o.ol1 = bld(bf.nl1, boff, soff, R(0));
        o.ol2 = bld(bf.nl2, boff, soff, R(0));
        o.cl0 = bld(bf.cl0, boff, soff, R(0));
        o.cl1 = bld(bf.cl1, boff, soff, R(0));
        o.cl2 = bld(bf.cl2, boff, soff, R(0));
    } else {
        o.ol0 = o.ol1 = o.ol2 = o.cl0 = o.cl1 = o.cl2 = R(0);
    }
    return o;
}

// Per-plane x data staged in LDS: hs = staggered spacing (sx[p-1], 1 off
// its range), hf = full spacing (fx[p], 1 off its range), dp = sponge term.
template <class R>
struct XPlane {
    R hs, hf, dp, pad;
};

template <class R> PF_DEV R sel(bool c, R a, R b) { return c ? a : b; }

// Own-cell store through the array's buffer descriptor, then, for the top
// cell of the column (k = Lz - 1), the z = max Mur face cell above it:
// face = mur(cur, s0, cur above), the cell above being the next lane.
template <int Q, int MODE, class R>
PF_DEV void put_tile(const Mesh<R>& m, const Bufs& bf, unsigned boff, unsigned soff, bool top, V<R> s0, V<R> cur,
                     V<R> above) {
    const Rsrc nh = Q == 0 ? bf.n0 : (Q == 1 ? bf.n1 : bf.n2);
    const Rsrc nl = Q == 0 ? bf.nl0 : (Q == 1 ? bf.nl1 : bf.nl2);
    bst(nh, boff, soff, s0.h);
    if constexpr (MODE == MODE_FP32C) bst(nl, boff, soff, s0.l);
    if (top) {
        const V<R> f = mur_map<MODE>(cur, m.mur[9 + Q], m.mur_lo[9 + Q], s0, above);
        bst(nh, boff + (unsigned)sizeof(R), soff, f.h);
        if constexpr (MODE == MODE_FP32C) bst(nl, boff + (unsigned)sizeof(R), soff, f.l);
    }
}

template <class R, int MODE, bool SPONGE, int TJ>
__global__ __launch_bounds__(TK * TJ, (MODE == MODE_FP32 ? 4 : 2)) void k_sweep_tiled(Mesh<R> m, TileMap tm) {
    constexpr bool EXACT = MODE == MODE_FP64;
    constexpr bool COMP = MODE == MODE_FP32C;
    __shared__ R su[2][3][TJ][TK];   // u planes, by plane parity
    __shared__ R sn2[2][TJ][TK];     // T2 of a plane, by plane parity
    __shared__ R ss4[TJ][TK];        // T4 of the current plane
    __shared__ R ss6[TJ][TK];        // T6 of the current plane
    __shared__ uint8_t sid[2][TJ][TK];
    __shared__ R smat[PF_TILED_MAX_MATERIALS * MAT_STRIDE];
    __shared__ XPlane<R> sxp[XT];    // planes ib-1 .. ie+1

    const int lane = threadIdx.x, row = threadIdx.y;
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int kt = t % tm.nkt;
    const int r_ = t / tm.nkt;
    const int jt = r_ % tm.njt;
    const int xc = r_ / tm.njt;
    const int k = kt * KOUT - 1 + lane;
    const int j = jt * (TJ - 2) - 1 + row;  // uniform over the wavefront
    const int ib = xc * tm.xchunk;
    const int ie = min(ib + tm.xchunk, m.nx);
    const bool kval = k >= 0 && k < m.nz;
    const bool jval = j >= 0 && j < m.ny;
    const bool valid = kval && jval;
    const bool writer = valid && lane >= 1 && lane <= TK - 2 && row >= 1 && row <= TJ - 2;
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;  // uniform over the grid
    const R w = (R)m.wave[tt];
    const int rp = row + 1 < TJ ? row + 1 : TJ - 1;
    const int rm = row > 0 ? row - 1 : 0;
    const int tid = row * TK + lane;

    for (int e = tid; e < m.nmat * MAT_STRIDE; e += TK * TJ) smat[e] = m.mat[e];
    for (int e = tid; e < ie - ib + 3; e += TK * TJ) {
        const int p = ib - 1 + e;
        XPlane<R> x;
        x.hs = (p >= 1 && p <= m.nx - 2) ? m.sx[p - 1] : R(1);
        x.hf = (p >= 0 && p <= m.nx - 2) ? m.fx[p] : R(1);
        x.dp = (SPONGE && p >= 0 && p < m.nx) ? m.dpx[p] : R(0);
        x.pad = R(0);
        sxp[e] = x;
    }

    // Column constants.  Every k = 0 (free surface) variant of the
    // reference's expressions is the interior expression with other
    // spacings / operands, selected here once: the per-value formulas are
    // then evaluated without branches and give the same bits as the
    // separate surface expressions (x/1 = x and x - 0 = x exactly).
    const bool ks = k >= 1;                    // interior in z
    const bool jn = j >= 1 && j <= m.ny - 2;   // staggered y index valid
    const bool jf = j >= 0 && j <= m.ny - 2;   // full y index valid
    const bool kn = k >= 1 && k <= m.nz - 2;
    const bool kf = k >= 0 && k <= m.nz - 2;
    const R hsy = jn ? m.sy[j - 1] : R(1);
    const R hfy = jf ? m.fy[j] : R(1);
    const R hsz = kn ? m.sz[k - 1] : R(1);
    const R hfz = kf ? m.fz[k] : R(1);
    const R hy_n = sel(ks, hsy, m.sy0);        // normal stresses, ux update: y
    const R hz_n = sel(ks, hsz, m.sz0);        // normal stresses, ux/uy update: z
    const R s4_h1 = sel(ks, hfz, m.fy0), s4_h2 = sel(ks, hfy, m.fz0);
    const R s5_h1 = sel(ks, hfz, m.fx0);
    const R s6_h1 = sel(ks, hfy, m.fx0);
    const R uy_g2 = sel(ks, hfy, m.fy0);
    const R uz_g3 = sel(ks, hfz, R(1));
    R dpyz = R(0), dpz_k = R(0);
    if constexpr (SPONGE) {
        dpyz = m.dpy[j < 0 ? 0 : (j < m.ny ? j : m.ny - 1)];
        dpz_k = m.dpz[k < 0 ? 0 : (k < m.nz ? k : m.nz - 1)];
    }
    const unsigned col = valid ? (unsigned)(j * m.pz + k) : 0u;
    const unsigned boff = col * (unsigned)sizeof(R);
    const unsigned plane = (unsigned)m.plane;
    // non-face cells per component (i-independent part); top = z-face below
    const bool wx = writer && jn && k <= m.nz - 2;
    const bool wy = writer && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2;
    const bool wz = writer && jn && k <= m.nz - 3;
    const bool top_xy = k == m.nz - 2;
    const bool top_z = k == m.nz - 3;

    const Bufs bf = make_bufs(m);
    // prologue: plane ib-1 (warm-up, previous-plane buffer 1), then the
    // loads in flight for the first plane of each parity
    // Loads of planes the chunk does not need are clamped to plane ie (or
    // nx-1), which the block reads anyway; planes outside the mesh are then
    // zeroed by settle().
    const int pmax = ie < m.nx ? ie : m.nx - 1;
    Col<R> c0 = settle(load_raw<R>(bf, ib >= 1 ? ib - 1 : 0, plane, boff, col), valid && ib >= 1,
                       ib - 1 == 0 && k == 0, w);
    // issue order ra, rb, ob (kept by the scheduling barriers) matches the
    // steady state, so the loop needs no wait for more than it uses
    Raw<R> ra = load_raw<R>(bf, ib, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Raw<R> rb = load_raw<R>(bf, ib + 1 <= pmax ? ib + 1 : pmax, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> ob = load_old<R, COMP>(bf, ib, plane, boff);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> oa = {};  // the warm-up plane loads plane ib+1 into it
    su[1][0][row][lane] = c0.x;
    su[1][1][row][lane] = c0.y;
    su[1][2][row][lane] = c0.z;
    sid[1][row][lane] = (uint8_t)c0.id;

    R n1_0 = 0, n2_0 = 0, n3_0 = 0;  // N at plane i
    R s5_m = 0, s6_m = 0;            // T5, T6 at plane i-1

    // One plane i: H is the LDS parity of plane i+1, rin holds the raw
    // column of plane i+1 and receives plane i+3, oin holds the node values
    // of plane i and receives plane i+2.
    auto plane_step = [&](int i, auto Hc, Raw<R>& rin, Old<R>& oin) {
        constexpr int H = decltype(Hc)::value;
        const int ip = i + 1;
        const Col<R> c1 = settle(rin, valid && ip < m.nx, ip == 0 && k == 0, w);
        su[H][0][row][lane] = c1.x;
        su[H][1][row][lane] = c1.y;
        su[H][2][row][lane] = c1.z;
        sid[H][row][lane] = (uint8_t)c1.id;
        __syncthreads();
        // issued after the barrier, so the registers of rin are free again
        rin = load_raw<R>(bf, ip + 2 <= pmax ? ip + 2 : pmax, plane, boff, col);

        const XPlane<R> xq0 = sxp[i - ib + 1];  // x data of plane i
        const XPlane<R> xq1 = sxp[ip - ib + 1];
        const R uy1_jm = su[H][1][rm][lane];
        const R uz1_km = lane_below(c1.z);
        const R uy0_kp = lane_above(c0.y);
        const R ux0_kp = lane_above(c0.x);
        const R uz0_jp = su[1 - H][2][rp][lane];
        const R ux0_jp = su[1 - H][0][rp][lane];
        const int id_jp = sid[1 - H][rp][lane];     // (i, j+1, k)
        const int id1_jp = sid[H][rp][lane];        // (i+1, j+1, k)
        const int id_kp = dpp_from_above(c0.id);    // (i, j, k+1)
        const int id_jkp = dpp_from_above(id_jp);   // (i, j+1, k+1)
        const int id1_kp = dpp_from_above(c1.id);   // (i+1, j, k+1)

        // normal stresses at plane ip (base_solver.py:329-351, 408-418)
        R n1_1, n2_1, n3_1;
        {
            const R* c = smat + MAT_STRIDE * c1.id;
            normal3<EXACT>(c, c1.x - c0.x, c1.y - uy1_jm, c1.z - sel(ks, uz1_km, R(0)), sel(ks, xq1.hs, m.sx0), hy_n,
                           hz_n, n1_1, n2_1, n3_1);
            const bool nv = ip >= 1 && ip <= m.nx - 2 && jn && kf;
            n1_1 = sel(nv, n1_1, R(0));
            n2_1 = sel(nv, n2_1, R(0));
            n3_1 = sel(nv && ks, n3_1, R(0));
        }
        // shear stresses at plane i (base_solver.py:353-372, 420-433)
        const bool ii = i >= 0 && i <= m.nx - 2;
        const R fxi = sel(ks, xq0.hf, m.fz0);
        R s4 = shear<EXACT>(smat[MAT_STRIDE * (ks ? id_jkp : id_jp) + M_C33], uy0_kp - c0.y, s4_h1,
                            uz0_jp - c0.z, s4_h2);
        R s5 = shear<EXACT>(smat[MAT_STRIDE * (ks ? id1_kp : c1.id) + M_C44], ux0_kp - c0.x, s5_h1,
                            c1.z - c0.z, fxi);
        R s6 = shear<EXACT>(smat[MAT_STRIDE * id1_jp + M_C55], ux0_jp - c0.x, s6_h1, c1.y - c0.y, fxi);
        s4 = sel(ii && i >= 1 && jf && kf, s4, R(0));
        s5 = sel(ii && jn && kf, s5, R(0));
        s6 = sel(ii && jf && kf, s6, R(0));
        sn2[H][row][lane] = n2_1;
        ss4[row][lane] = s4;
        ss6[row][lane] = s6;
        __syncthreads();

        const R n3_0_kp = lane_above(n3_0);
        const R s4_km = lane_below(s4);
        const R s5_km = lane_below(s5);
        if (i >= ib && i < ie) {
            const R n2_0_jp = sn2[1 - H][rp][lane];
            const R s4_jm = ss4[rm][lane];
            const R s6_jm = ss6[rm][lane];
            const R s = SPONGE ? (xq0.dp + dpyz) + dpz_k : R(0);
            const R gx = sel(ks, xq0.hs, m.sx0);
            const unsigned soff = uni((unsigned)i * plane * (unsigned)sizeof(R));
            const R cl2 = (i == 0 && k == 0) ? R(0) : oin.cl2;
            // ux (base_solver.py:441-447, 495-501)
            {
                const R dv = div3<EXACT>(n1_1 - n1_0, sel(ks, xq0.hf, m.fx0), s6 - s6_jm, hy_n,
                                         s5 - sel(ks, s5_km, R(0)), hz_n);
                const V<R> cur = {c0.x, oin.cl0};
                const V<R> s0 =
                    leap_from<MODE, SPONGE>(cur, oin.o0, oin.ol0, smat[MAT_STRIDE * c1.id + M_COEF], dv, s);
                const V<R> above = {ux0_kp, COMP ? lane_above(oin.cl0) : R(0)};
                if (wx && i <= m.nx - 3) put_tile<0, MODE>(m, bf, boff, soff, top_xy, s0, cur, above);
            }
            // uy (base_solver.py:449-455, 503-509)
            {
                const R dv = div3<EXACT>(s6 - s6_m, gx, n2_0_jp - n2_0, uy_g2, s4 - sel(ks, s4_km, R(0)), hz_n);
                const V<R> cur = {c0.y, oin.cl1};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, oin.o1, oin.ol1, smat[MAT_STRIDE * id_jp + M_COEF], dv, s);
                if (i == 0) s0 = {R(0), R(0)};
                const V<R> above = {uy0_kp, COMP ? lane_above(oin.cl1) : R(0)};
                if (wy && i <= m.nx - 2) put_tile<1, MODE>(m, bf, boff, soff, top_xy, s0, cur, above);
            }
            // uz (base_solver.py:457-463, 511-517)
            {
                const R dv = div3<EXACT>(s5 - s5_m, gx, s4 - s4_jm, hy_n, n3_0_kp - n3_0, uz_g3);
                const V<R> cur = {c0.z, cl2};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, oin.o2, oin.ol2,
                                                  smat[MAT_STRIDE * (ks ? id_kp : c0.id) + M_COEF], dv, s);
                if (i == 0) s0 = {R(0), R(0)};
                const V<R> above = {lane_above(c0.z), COMP ? lane_above(cl2) : R(0)};
                if (wz && i <= m.nx - 2) put_tile<2, MODE>(m, bf, boff, soff, top_z, s0, cur, above);
            }
        }
        oin = load_old<R, COMP>(bf, i + 2 < ie ? i + 2 : ie - 1, plane, boff);
        c0 = c1;
        n1_0 = n1_1;
        n2_0 = n2_1;
        n3_0 = n3_1;
        s5_m = s5;
        s6_m = s6;
    };

    // planes ib-1, ib+1, ... use parity 0 for plane i+1; ib, ib+2, ... parity 1
    // (an odd count of planes ends with a parity-1 pass over plane ie that
    // only loads and computes: every path through the loop body issues the
    // same loads, so no wait is made for the shortest of several paths)
    for (int i = ib - 1; i < ie; i += 2) {
        plane_step(i, std::integral_constant<int, 0>(), ra, oa);
        plane_step(i + 1, std::integral_constant<int, 1>(), rb, ob);
    }
}

inline TileMap tile_map(const pf_ctx* c, int tj, int xchunk) {
    TileMap tm;
    tm.nkt = (c->nz + KOUT - 1) / KOUT;
    tm.njt = (c->ny + (tj - 2) - 1) / (tj - 2);
    if (xchunk <= 0) {
        // aim for >= ~2048 blocks (8 per CU) with chunks of at least 16 planes
        const long long per = (long long)tm.nkt * tm.njt;
        long long want = (2048 + per - 1) / per;
        if (want < 1) want = 1;
        xchunk = (int)((c->nx + want - 1) / want);
        if (xchunk < 16) xchunk = 16;
        if (xchunk > XT - 3) xchunk = XT - 3;
    }
    if (xchunk > c->nx) xchunk = c->nx;
    tm.xchunk = xchunk;
    tm.nxc = (c->nx + xchunk - 1) / xchunk;
    return tm;
}

template <class R, int MODE, bool SPONGE>
int launch_tiled_s(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    const int tj = c->tile_j;
    if ((unsigned long long)c->cells * sizeof(R) >= (1ULL << 32))
        return fail(PF_ERR_ARG, "tiled sweep: field arrays must be < 4 GiB (use the naive kernel)");
    if (c->tile_x > XT - 3) return fail(PF_ERR_ARG, "tile_x must be <= %d (got %d)", XT - 3, c->tile_x);
    TileMap tm = tile_map(c, tj, c->tile_x);
    const long long nb = (long long)tm.nkt * tm.njt * tm.nxc;
    if (nb > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many tiles");
    const dim3 grd((unsigned)nb);
    switch (tj) {
        case 4: hipLaunchKernelGGL((k_sweep_tiled<R, MODE, SPONGE, 4>), grd, dim3(TK, 4), 0, s, m, tm); break;
        case 8: hipLaunchKernelGGL((k_sweep_tiled<R, MODE, SPONGE, 8>), grd, dim3(TK, 8), 0, s, m, tm); break;
        case 16: hipLaunchKernelGGL((k_sweep_tiled<R, MODE, SPONGE, 16>), grd, dim3(TK, 16), 0, s, m, tm); break;
        default: return fail(PF_ERR_ARG, "tile_j must be 4, 8 or 16 (got %d)", tj);
    }
    HIPCHK(hipGetLastError());
    return PF_OK;
}

template <class R, int MODE>
int launch_tiled(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    return c->sponge ? launch_tiled_s<R, MODE, true>(c, m, s) : launch_tiled_s<R, MODE, false>(c, m, s);
}

}  // namespace pf
