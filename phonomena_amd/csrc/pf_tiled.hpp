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

// Plane accesses go through buffer resources: one SGPR descriptor per time
// level (whose arrays ux, uy, uz (, lo parts) lie back to back) and one for
// the material indices, the array and plane selected by the scalar byte
// offset q * cells * sizeof(R) + ip * plane * sizeof(R) and the cell by a
// per-lane 32-bit byte offset that does not change along the march, so a
// load or store costs no per-lane address arithmetic.  A level must be
// smaller than 4 GiB (the launcher checks); every access made here lies
// inside its level.
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
PF_DEV lo_t bld_lo(Rsrc r, unsigned off, unsigned soff) {
    return (lo_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, (int)soff, 0);
}
PF_DEV void bst_lo(Rsrc r, unsigned off, unsigned soff, lo_t v) {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)off, (int)soff, 0);
}

// A block-uniform value kept in a scalar register (plane offsets).
PF_DEV unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// Descriptors of the two levels and the index map.  A level holds the three
// hi arrays (ab bytes each) and, in FP32C, three 16-bit lo arrays (lb bytes
// each) from byte 3 * ab on.
struct Bufs {
    Rsrc cur, old, nxt, ids;  // level n (read), n-1 (read), n+1 (written)
    unsigned ab, lb;
    // Lane offset past the end of a level (0: none): a buffer load there
    // returns 0 without reading memory, which lanes whose value is never
    // used take.  Only when a level is < 2 GiB, so no offset sum wraps.
    unsigned far;
};
template <class R, bool COMP> PF_DEV Bufs make_bufs(const Mesh<R>& m) {
    const long long n = (long long)m.nx * m.plane;
    const long long lvl = n * 3 * (long long)sizeof(R) + (COMP ? n * 3 * (long long)sizeof(lo_t) : 0);
    Bufs b;
    b.cur = __builtin_amdgcn_make_buffer_rsrc(const_cast<R*>(m.c0), (short)0, (int)(unsigned)lvl, PF_RSRC_FLAGS);
    b.old = __builtin_amdgcn_make_buffer_rsrc(const_cast<R*>(m.o0), (short)0, (int)(unsigned)lvl, PF_RSRC_FLAGS);
    b.nxt = __builtin_amdgcn_make_buffer_rsrc(m.n0, (short)0, (int)(unsigned)lvl, PF_RSRC_FLAGS);
    b.ids = arr_rsrc(m.ids, n);
    b.ab = uni((unsigned)(n * (long long)sizeof(R)));
    b.lb = uni((unsigned)(n * (long long)sizeof(lo_t)));
    b.far = lvl < (1LL << 31) ? 0x80000000u : 0u;
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
    r.x = bld(bf.cur, boff, soff, R(0));
    r.y = bld(bf.cur, boff, soff + bf.ab, R(0));
    r.z = bld(bf.cur, boff, soff + 2 * bf.ab, R(0));
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
// previous level (hi, and the stored 16-bit lo in FP32C) and the current
// level's stored lo part (decoded at use, against their hi).  The previous
// level is read at lane offsets ooff / oloff (the far offset for lanes that
// do not use it).
template <class R>
struct Old {
    R o0, o1, o2;
    lo_t ol0, ol1, ol2, cl0, cl1, cl2;
};

template <class R, bool COMP>
PF_DEV Old<R> load_old(const Bufs& bf, int i, unsigned plane, unsigned boff, unsigned loff, unsigned ooff,
                       unsigned oloff) {
    const unsigned soff = uni((unsigned)i * plane * (unsigned)sizeof(R));
    Old<R> o;
    o.o0 = bld(bf.old, ooff, soff, R(0));
    o.o1 = bld(bf.old, ooff, soff + bf.ab, R(0));
    o.o2 = bld(bf.old, ooff, soff + 2 * bf.ab, R(0));
    if constexpr (COMP) {
        const unsigned sl = uni((unsigned)i * plane * (unsigned)sizeof(lo_t)) + 3 * bf.ab;
        o.ol0 = bld_lo(bf.old, oloff, sl);
        o.ol1 = bld_lo(bf.old, oloff, sl + bf.lb);
        o.ol2 = bld_lo(bf.old, oloff, sl + 2 * bf.lb);
        o.cl0 = bld_lo(bf.cur, loff, sl);
        o.cl1 = bld_lo(bf.cur, loff, sl + bf.lb);
        o.cl2 = bld_lo(bf.cur, loff, sl + 2 * bf.lb);
    } else {
        o.ol0 = o.ol1 = o.ol2 = o.cl0 = o.cl1 = o.cl2 = 0;
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
PF_DEV void put_tile(const Mesh<R>& m, const Bufs& bf, unsigned boff, unsigned loff, unsigned soff, unsigned slo,
                     bool top, V<R> s0, V<R> cur, V<R> above) {
    const unsigned sh = soff + Q * bf.ab, sl = slo + 3 * bf.ab + Q * bf.lb;
    bst(bf.nxt, boff, sh, s0.h);
    if constexpr (MODE == MODE_FP32C) bst_lo(bf.nxt, loff, sl, lo_encode(s0.h, s0.l));
    if (top) {
        const V<R> f = mur_map<MODE>(cur, m.mur[9 + Q], m.mur_lo[9 + Q], s0, above);
        bst(bf.nxt, boff + (unsigned)sizeof(R), sh, f.h);
        if constexpr (MODE == MODE_FP32C) bst_lo(bf.nxt, loff + (unsigned)sizeof(lo_t), sl, lo_encode(f.h, f.l));
    }
}

// Shared memory of one block.  Material rows are addressed by byte offset
// (id * row bytes), which is what the id registers and sid carry.
template <class R, int TJ, int KZ>
struct TiledLds {
    R su[2][3][TJ][KZ];   // u planes, by plane parity
    R sn2[2][TJ][KZ];     // T2 of a plane, by plane parity
    R ss4[TJ][KZ];        // T4 of the current plane
    R ss6[TJ][KZ];        // T6 of the current plane
    uint16_t sid[2][TJ][KZ];
    R smat[PF_TILED_MAX_MATERIALS * MAT_STRIDE];
    XPlane<R> sxp[XT];    // planes ib-1 .. ie+1
};

template <class R> PF_DEV const R* mrow(const R* smat, int off) {
    return (const R*)((const char*)smat + off);
}
template <class R> PF_DEV R mget(const R* smat, int off, int e) { return mrow(smat, off)[e]; }

// DPP shifts filling 0 where the wavefront has no source lane (one
// instruction; only halo lanes see the fill).
PF_DEV int dpp_below0(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
PF_DEV int dpp_above0(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
PF_DEV float below0(float v) { return __int_as_float(dpp_below0(__float_as_int(v))); }
PF_DEV float above0(float v) { return __int_as_float(dpp_above0(__float_as_int(v))); }
PF_DEV double below0(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_below0((int)(b & 0xffffffffLL)), hi = dpp_below0((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
PF_DEV double above0(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_above0((int)(b & 0xffffffffLL)), hi = dpp_above0((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The march of one block over its tile and chunk.  FYZ: every lane of the
// tile, halos included, lies at 1 <= k <= nz-4 and 1 <= j <= ny-3, so no
// lane meets the free surface, a y or z range limit or face; FX: the chunk
// lies at 2 <= ib, ie <= nx-3, clear of the x = 0 wall and the x face.
// Where a flag holds, the range selects it covers are the identity and are
// left out (same bits); the x conditions left in a FYZ block are uniform
// over the block (per plane).
template <class R, int MODE, bool SPONGE, int TJ, int KZ, bool FYZ, bool FX>
PF_DEV void tiled_march(const Mesh<R>& m, TiledLds<R, TJ, KZ>& L, int kl, int row, int k, int j, int ib, int ie,
                        R w) {
    constexpr int EXACT = exact_of(MODE);
    constexpr bool COMP = MODE == MODE_FP32C;
    constexpr int RB = MAT_STRIDE * (int)sizeof(R);  // bytes per material row
    const int rp = row + 1 < TJ ? row + 1 : TJ - 1;
    const int rm = row > 0 ? row - 1 : 0;

    // Column constants.  Every k = 0 (free surface) variant of the
    // reference's expressions is the interior expression with other
    // spacings / operands, selected here once: the per-value formulas are
    // then evaluated without branches and give the same bits as the
    // separate surface expressions (x/1 = x and x - 0 = x exactly).
    constexpr bool FULL = FYZ && FX;
    const bool kval = FYZ || (k >= 0 && k < m.nz);
    const bool jval = FYZ || (j >= 0 && j < m.ny);
    const bool valid = kval && jval;
    const bool ks = FYZ || k >= 1;                          // interior in z
    const bool jn = FYZ || (j >= 1 && j <= m.ny - 2);       // staggered y index valid
    const bool jf = FYZ || (j >= 0 && j <= m.ny - 2);       // full y index valid
    const bool kn = FYZ || (k >= 1 && k <= m.nz - 2);
    const bool kf = FYZ || (k >= 0 && k <= m.nz - 2);
    const bool writer = valid && kl >= 1 && kl <= KZ - 2 && row >= 1 && row <= TJ - 2;
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
    // u_z's T3 term at k = 0: the reference's unscaled T3[..,1], or (surf)
    // the difference divided by fdz, as in the interior
    const R uz_g3 = sel(ks || m.surf, hfz, R(1));
    R dpyz = R(0), dpz_k = R(0);
    if constexpr (SPONGE) {
        dpyz = m.dpy[j < 0 ? 0 : (j < m.ny ? j : m.ny - 1)];
        dpz_k = m.dpz[k < 0 ? 0 : (k < m.nz ? k : m.nz - 1)];
    }
    const unsigned col = valid ? (unsigned)(j * m.pz + k) : 0u;
    const unsigned boff = col * (unsigned)sizeof(R);
    const unsigned loff = col * (unsigned)sizeof(lo_t);
    const unsigned plane = (unsigned)m.plane;
    // non-face cells per component (i-independent part); top = z-face below
    const bool wx = writer && jn && k <= m.nz - 2;
    const bool wy = writer && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2;
    const bool wz = writer && jn && k <= m.nz - 3;
    const bool top_xy = !FYZ && k == m.nz - 2;
    const bool top_z = !FYZ && k == m.nz - 3;

    const Bufs bf = make_bufs<R, COMP>(m);
    // only writer lanes use the previous level
    const unsigned ooff = writer || !bf.far ? boff : bf.far;
    const unsigned oloff = writer || !bf.far ? loff : bf.far;
    // prologue: plane ib-1 (warm-up, previous-plane buffer 1), then the
    // loads in flight for the first plane of each parity.  Loads of planes
    // the chunk does not need are clamped to plane ie (or nx-1), which the
    // block reads anyway; planes outside the mesh are then zeroed by settle().
    const int pmax = ie < m.nx ? ie : m.nx - 1;
    Col<R> c0 = settle(load_raw<R>(bf, ib >= 1 ? ib - 1 : 0, plane, boff, col), valid && ib >= 1,
                       !FYZ && ib - 1 == 0 && k == 0, w);
    c0.id *= RB;
    // issue order ra, rb, ob (kept by the scheduling barriers) matches the
    // steady state, so the loop needs no wait for more than it uses
    Raw<R> ra = load_raw<R>(bf, ib, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Raw<R> rb = load_raw<R>(bf, ib + 1 <= pmax ? ib + 1 : pmax, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> ob = load_old<R, COMP>(bf, ib, plane, boff, loff, ooff, oloff);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> oa = {};  // the warm-up plane loads plane ib+1 into it
    L.su[1][0][row][kl] = c0.x;
    L.su[1][1][row][kl] = c0.y;
    L.su[1][2][row][kl] = c0.z;
    L.sid[1][row][kl] = (uint16_t)c0.id;

    R n1_0 = 0, n2_0 = 0, n3_0 = 0;  // N at plane i
    R s5_m = 0, s6_m = 0;            // T5, T6 at plane i-1
    const R* smat = L.smat;

    // One plane i: H is the LDS parity of plane i+1, rin holds the raw
    // column of plane i+1 and receives plane i+3, oin holds the node values
    // of plane i and receives plane i+2.
    auto plane_step = [&](int i, auto Hc, Raw<R>& rin, Old<R>& oin) {
        constexpr int H = decltype(Hc)::value;
        const int ip = i + 1;
        Col<R> c1 = settle(rin, valid && (FX || ip < m.nx), !FYZ && ip == 0 && k == 0, w);
        c1.id *= RB;
        L.su[H][0][row][kl] = c1.x;
        L.su[H][1][row][kl] = c1.y;
        L.su[H][2][row][kl] = c1.z;
        L.sid[H][row][kl] = (uint16_t)c1.id;
        __syncthreads();
        // issued after the barrier, so the registers of rin are free again
        rin = load_raw<R>(bf, ip + 2 <= pmax ? ip + 2 : pmax, plane, boff, col);

        const XPlane<R> xq0 = L.sxp[i - ib + 1];  // x data of plane i
        const XPlane<R> xq1 = L.sxp[ip - ib + 1];
        const R uy1_jm = L.su[H][1][rm][kl];
        const R uz1_km = below0(c1.z);
        const R uy0_kp = above0(c0.y);
        const R ux0_kp = above0(c0.x);
        const R uz0_jp = L.su[1 - H][2][rp][kl];
        const R ux0_jp = L.su[1 - H][0][rp][kl];
        const int id_jp = L.sid[1 - H][rp][kl];   // (i, j+1, k)
        const int id1_jp = L.sid[H][rp][kl];      // (i+1, j+1, k)
        const int id_kp = dpp_above0(c0.id);      // (i, j, k+1)
        const int id_jkp = dpp_above0(id_jp);     // (i, j+1, k+1)
        const int id1_kp = dpp_above0(c1.id);     // (i+1, j, k+1)

        // normal stresses at plane ip (base_solver.py:329-351, 408-418)
        R n1_1, n2_1, n3_1;
        {
            normal3<EXACT>(mrow(smat, c1.id), c1.x - c0.x, c1.y - uy1_jm, c1.z - sel(ks, uz1_km, R(0)),
                           sel(ks, xq1.hs, m.sx0), hy_n, hz_n, n1_1, n2_1, n3_1);
            if constexpr (!FULL) {
                const bool nv = ip >= 1 && ip <= m.nx - 2 && jn && kf;
                n1_1 = sel(nv, n1_1, R(0));
                n2_1 = sel(nv, n2_1, R(0));
                n3_1 = sel(nv && ks, n3_1, R(0));
            }
        }
        // shear stresses at plane i (base_solver.py:353-372, 420-433)
        const R fxi = sel(ks, xq0.hf, m.fz0);
        R s4 = shear<EXACT>(mget(smat, ks ? id_jkp : id_jp, M_C33), uy0_kp - c0.y, s4_h1, uz0_jp - c0.z, s4_h2);
        R s5 = shear<EXACT>(mget(smat, ks ? id1_kp : c1.id, M_C44), ux0_kp - c0.x, s5_h1, c1.z - c0.z, fxi);
        R s6 = shear<EXACT>(mget(smat, id1_jp, M_C55), ux0_jp - c0.x, s6_h1, c1.y - c0.y, fxi);
        if constexpr (!FULL) {
            const bool ii = i >= 0 && i <= m.nx - 2;
            s4 = sel(ii && i >= 1 && jf && kf, s4, R(0));
            s5 = sel(ii && jn && kf, s5, R(0));
            s6 = sel(ii && jf && kf, s6, R(0));
        }
        L.sn2[H][row][kl] = n2_1;
        L.ss4[row][kl] = s4;
        L.ss6[row][kl] = s6;
        __syncthreads();

        const R n3_0_kp = above0(n3_0);
        const R s4_km = below0(s4);
        const R s5_km = below0(s5);
        if (i >= ib && i < ie) {
            const R n2_0_jp = L.sn2[1 - H][rp][kl];
            const R s4_jm = L.ss4[rm][kl];
            const R s6_jm = L.ss6[rm][kl];
            const R s = SPONGE ? (xq0.dp + dpyz) + dpz_k : R(0);
            const R gx = sel(ks, xq0.hs, m.sx0);
            const unsigned soff = uni((unsigned)i * plane * (unsigned)sizeof(R));
            const unsigned slo = uni((unsigned)i * plane * (unsigned)sizeof(lo_t));
            // lo parts decoded against their stored hi (c0 is the stored
            // value except at the wall source line, whose lo is 0)
            const R cl0 = COMP ? lo_decode(c0.x, oin.cl0) : R(0);
            const R cl1 = COMP ? lo_decode(c0.y, oin.cl1) : R(0);
            const R cl2 = (!COMP || (!FYZ && i == 0 && k == 0)) ? R(0) : lo_decode(c0.z, oin.cl2);
            const R ol0 = COMP ? lo_decode(oin.o0, oin.ol0) : R(0);
            const R ol1 = COMP ? lo_decode(oin.o1, oin.ol1) : R(0);
            const R ol2 = COMP ? lo_decode(oin.o2, oin.ol2) : R(0);
            const bool wall = !FX && i == 0;  // uy, uz on the x = 0 wall stay 0
            // ux (base_solver.py:441-447, 495-501)
            {
                const R dv = div3<EXACT>(n1_1 - n1_0, sel(ks, xq0.hf, m.fx0), s6 - s6_jm, hy_n,
                                         s5 - sel(ks, s5_km, R(0)), hz_n);
                const V<R> cur = {c0.x, cl0};
                const V<R> s0 = leap_from<MODE, SPONGE>(cur, oin.o0, ol0, mget(smat, c1.id, M_COEF), dv, s);
                const V<R> above = {ux0_kp, COMP ? above0(cl0) : R(0)};
                if (wx && (FX || i <= m.nx - 3))
                    put_tile<0, MODE>(m, bf, boff, loff, soff, slo, top_xy, s0, cur, above);
            }
            // uy (base_solver.py:449-455, 503-509)
            {
                const R dv = div3<EXACT>(s6 - s6_m, gx, n2_0_jp - n2_0, uy_g2, s4 - sel(ks, s4_km, R(0)), hz_n);
                const V<R> cur = {c0.y, cl1};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, oin.o1, ol1, mget(smat, id_jp, M_COEF), dv, s);
                if (wall) s0 = {R(0), R(0)};
                const V<R> above = {uy0_kp, COMP ? above0(cl1) : R(0)};
                if (wy && (FX || i <= m.nx - 2))
                    put_tile<1, MODE>(m, bf, boff, loff, soff, slo, top_xy, s0, cur, above);
            }
            // uz (base_solver.py:457-463, 511-517)
            {
                const R dv = div3<EXACT>(s5 - s5_m, gx, s4 - s4_jm, hy_n, n3_0_kp - n3_0, uz_g3);
                const V<R> cur = {c0.z, cl2};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, oin.o2, ol2, mget(smat, ks ? id_kp : c0.id, M_COEF), dv, s);
                if (wall) s0 = {R(0), R(0)};
                const V<R> above = {above0(c0.z), COMP ? above0(cl2) : R(0)};
                if (wz && (FX || i <= m.nx - 2))
                    put_tile<2, MODE>(m, bf, boff, loff, soff, slo, top_z, s0, cur, above);
            }
        }
        oin = load_old<R, COMP>(bf, i + 2 < ie ? i + 2 : ie - 1, plane, boff, loff, ooff, oloff);
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

// One tile of a sweep: tile index t (after any remapping) of mesh m under
// tile map tm, LDS L.  Tile row jt starts at row jt * jstep - 1 and x chunk
// xc at plane x0 + xc * xchunk (a sweep of the whole mesh: jstep TJ-2, x0
// 0).  Every thread of the block calls it; it ends without a barrier, so a
// caller that reuses L for another tile syncs first.
template <class R, int MODE, bool SPONGE, int TJ, int KZ>
PF_DEV void tiled_tile(const Mesh<R>& m, const TileMap& tm, int t, TiledLds<R, TJ, KZ>& L, int jstep = TJ - 2,
                       int x0 = 0) {
    const int kl = threadIdx.x % KZ;
    const int row = threadIdx.y * (64 / KZ) + threadIdx.x / KZ;
    const int kt = t % tm.nkt;
    const int r_ = t / tm.nkt;
    const int jt = r_ % tm.njt;
    const int xc = r_ / tm.njt;
    const int k = kt * (KZ - 2) - 1 + kl;
    const int j = jt * jstep - 1 + row;
    const int ib = x0 + xc * tm.xchunk;
    const int ie = min(ib + tm.xchunk, m.nx);
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;  // uniform over the grid
    const R w = (R)m.wave[tt];
    const int tid = threadIdx.y * 64 + threadIdx.x;

    for (int e = tid; e < m.nmat * MAT_STRIDE; e += TJ * KZ) L.smat[e] = m.mat[e];
    for (int e = tid; e < ie - ib + 3; e += TJ * KZ) {
        const int p = ib - 1 + e;
        XPlane<R> x;
        x.hs = (p >= 1 && p <= m.nx - 2) ? m.sx[p - 1] : R(1);
        x.hf = (p >= 0 && p <= m.nx - 2) ? m.fx[p] : R(1);
        x.dp = (SPONGE && p >= 0 && p < m.nx) ? m.dpx[p] : R(0);
        x.pad = R(0);
        L.sxp[e] = x;
    }
    // block-uniform: the whole tile and chunk are clear of every boundary
    const int k_lo = kt * (KZ - 2) - 1, k_hi = k_lo + KZ - 1;
    const int j_lo = jt * jstep - 1, j_hi = j_lo + TJ - 1;
    const bool fyz = k_lo >= 1 && k_hi <= m.nz - 4 && j_lo >= 1 && j_hi <= m.ny - 3;
    const bool fx = ib >= 2 && ie <= m.nx - 3;
    if (fyz && fx)
        tiled_march<R, MODE, SPONGE, TJ, KZ, true, true>(m, L, kl, row, k, j, ib, ie, w);
    else if (fyz)
        tiled_march<R, MODE, SPONGE, TJ, KZ, true, false>(m, L, kl, row, k, j, ib, ie, w);
    else
        tiled_march<R, MODE, SPONGE, TJ, KZ, false, false>(m, L, kl, row, k, j, ib, ie, w);
}

// KZ lanes along z per tile row (64: one row per wavefront; 32: two rows
// per wavefront, lanes 0-31 and 32-63), TJ rows along y per block.
template <class R, int MODE, bool SPONGE, int TJ, int KZ>
__global__ __launch_bounds__(TJ * KZ, (MODE == MODE_FP32 ? 4 : 2)) void k_sweep_tiled(Mesh<R> m, TileMap tm) {
    __shared__ TiledLds<R, TJ, KZ> L;
    tiled_tile<R, MODE, SPONGE, TJ, KZ>(m, tm, xcd_remap((int)blockIdx.x, (int)gridDim.x), L);
}

// Blocks of a kernel that one CU holds at once (1 when the runtime cannot
// say).
template <class K> int blocks_per_cu(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 1;
    }
    return n;
}

// Planes per x chunk, from the rounds in which the device runs the blocks.
// per tiles per plane, each block marching xc planes plus `warm` planes it
// only loads and computes; `slots` blocks resident at once.  Over the
// balanced chunkings (n chunks of ceil(nx / n) planes, lo <= planes <= cap)
// the one with the most useful planes per resident-block plane time:
//     per * nx / (slots * rounds * (xc + warm)),  rounds = ceil(per * n / slots)
// which charges a last round that leaves CUs idle and the warm-up planes
// of short chunks.  Ties go to more chunks.
inline int round_xchunk(long long per, int nx, int warm, int lo, int cap, long long slots) {
    int best = nx < cap ? nx : cap;
    double best_s = -1.0;
    for (int n = 1; n <= nx; n++) {
        const int xc = (nx + n - 1) / n;
        if (xc > cap) continue;
        if (xc < lo && n > 1) break;
        const int nc = (nx + xc - 1) / xc;
        if (nc != n) continue;  // same chunking as a smaller n
        const long long blocks = per * nc;
        const long long rounds = (blocks + slots - 1) / slots;
        const double sc = (double)per * nx / ((double)slots * (double)rounds * (double)(xc + warm));
        if (sc >= best_s) {
            best_s = sc;
            best = xc;
        }
    }
    return best;
}

// slots > 0: x chunks by round_xchunk for that many resident blocks;
// otherwise a fixed rule.  An explicit xchunk > 0 is kept (within range).
inline TileMap tile_map(const pf_ctx* c, int tj, int kz, int xchunk, long long slots = 0) {
    TileMap tm;
    tm.nkt = (c->nz + (kz - 2) - 1) / (kz - 2);
    tm.njt = (c->ny + (tj - 2) - 1) / (tj - 2);
    if (xchunk <= 0 && slots > 0) {
        xchunk = round_xchunk((long long)tm.nkt * tm.njt, c->nx, 1, 16, XT - 3, slots);
    } else if (xchunk <= 0) {
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

// Lane-slot count of a tiling (lanes x planes), used to pick the z layout.
inline double tiled_slots(const pf_ctx* c, int tj, int kz) {
    const TileMap tm = tile_map(c, tj, kz, c->tile_x);
    return (double)tm.nkt * kz * tm.njt * tj * tm.nxc * (tm.xchunk + 2);
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
int launch_tiled_k(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    static const int per_cu = blocks_per_cu(k_sweep_tiled<R, MODE, SPONGE, TJ, KZ>, TJ * KZ);
    TileMap tm = tile_map(c, TJ, KZ, c->tile_x, (long long)c->cus * per_cu);
    const long long nb = (long long)tm.nkt * tm.njt * tm.nxc;
    if (nb > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many tiles");
    hipLaunchKernelGGL((k_sweep_tiled<R, MODE, SPONGE, TJ, KZ>), dim3((unsigned)nb), dim3(64, TJ * KZ / 64), 0, s, m,
                       tm);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

template <class R, int MODE, bool SPONGE>
int launch_tiled_s(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    if (level_bytes(c->cells, c->precision) >= (1ULL << 32))
        return fail(PF_ERR_ARG, "tiled sweep: a time level must be < 4 GiB (use the naive kernel)");
    if (c->tile_x > XT - 3) return fail(PF_ERR_ARG, "tile_x must be <= %d (got %d)", XT - 3, c->tile_x);
    const int tj = c->tile_j, kz = c->tile_z;
    if (kz == 64) {
        switch (tj) {
            case 4: return launch_tiled_k<R, MODE, SPONGE, 4, 64>(c, m, s);
            case 8: return launch_tiled_k<R, MODE, SPONGE, 8, 64>(c, m, s);
            case 16: return launch_tiled_k<R, MODE, SPONGE, 16, 64>(c, m, s);
        }
    } else if (kz == 32) {
        switch (tj) {
            case 8: return launch_tiled_k<R, MODE, SPONGE, 8, 32>(c, m, s);
            case 16: return launch_tiled_k<R, MODE, SPONGE, 16, 32>(c, m, s);
            case 32: return launch_tiled_k<R, MODE, SPONGE, 32, 32>(c, m, s);
        }
    }
    return fail(PF_ERR_ARG, "tiled sweep: tile_j %d with tile_z %d is not a built shape", tj, kz);
}

template <class R, int MODE>
int launch_tiled(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    return c->sponge ? launch_tiled_s<R, MODE, true>(c, m, s) : launch_tiled_s<R, MODE, false>(c, m, s);
}

}  // namespace pf
