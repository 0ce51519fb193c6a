// LDS-tiled sweep: (y, z) tiles, register march along x.
//
// A block is 64 lanes along z (one wavefront per row) by TJ rows along y and
// marches through a chunk of x planes.  Tiles overlap by one column / row on
// each side: lanes 0 and 63 and rows 0 and TJ-1 compute the stresses their
// neighbours need but write nothing, so every thread runs the same code and
// no halo special-casing is needed.  Per plane each thread
//   1. loads its own column of u at plane i+1 (one coalesced load per
//      component) and its material index into registers and LDS;
//   2. computes the normal stresses N = (T1, T2, T3) at plane i+1 and the
//      shear stresses T4, T5, T6 at plane i from register values, LDS
//      neighbours along y and cross-lane neighbours along z;
//   3. publishes the stresses its y neighbours need through LDS;
//   4. forms the displacement update of plane i (stage I) and writes it,
//      plus the z = max Mur face cell above it when it is the top cell.
// Own-column values of the previous plane (u, N, T5, T6) stay in registers,
// so every u value is read from HBM once per step and written once.
//
// LDS hazards: su/sid of a plane parity are rewritten at the top of the
// iteration two planes later, before that iteration's first barrier, so
// they are only read before the current iteration's second barrier;
// sn2/ss4/ss6 are rewritten after a first barrier and may be read after it.
//
// The arithmetic is the shared per-value formulas of pf_formulas.hpp, fed
// the same operands as the naive kernel, so both sweeps produce identical
// bits.  A chunk starting at plane ib first runs one warm-up plane (ib-1)
// that only builds the register/LDS state.
#pragma once
#include "pf_runtime.hpp"

namespace pf {

constexpr int TK = 64;        // lanes along z
constexpr int KOUT = TK - 2;  // output columns per tile

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

template <class R> PF_DEV R from_lane_below(R v) { return __shfl_up(v, 1, 64); }   // value of lane-1 (z-1)
template <class R> PF_DEV R from_lane_above(R v) { return __shfl_down(v, 1, 64); } // value of lane+1 (z+1)

// Own column of one plane: u components and material index.
template <class R>
struct Col {
    R x, y, z;
    int id;
};

// Plane accesses go through buffer resources: one SGPR descriptor per array
// (built once per block), the plane selected by the scalar byte offset
// ip * plane * sizeof(R) and the cell by a per-lane 32-bit byte offset that
// does not change along the march, so a load or store costs no per-lane
// address arithmetic.  Arrays must be smaller than 4 GiB (the launcher
// checks); accesses beyond an array's range read 0 / are dropped.
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

template <class R>
PF_DEV Col<R> load_col(const Mesh<R>& m, const Bufs& bf, int ip, bool valid, unsigned col, int k, R w) {
    const bool pin = ip >= 0 && ip < m.nx;
    const bool in = valid && pin;
    const unsigned pc = pin ? (unsigned)ip : 0u;
    const unsigned boff = col * (unsigned)sizeof(R);
    const unsigned soff = pc * (unsigned)m.plane * (unsigned)sizeof(R);
    const R x = bld(bf.c0, boff, soff, R(0));
    const R y = bld(bf.c1, boff, soff, R(0));
    const R z = bld(bf.c2, boff, soff, R(0));
    const int id = bld_u8(bf.ids, col, pc * (unsigned)m.plane);
    Col<R> c;
    c.x = in ? x : R(0);
    c.y = in ? y : R(0);
    c.z = in ? ((ip == 0 && k == 0) ? w : z) : R(0);
    c.id = in ? id : 0;
    return c;
}

template <class R> PF_DEV R sel(bool c, R a, R b) { return c ? a : b; }

// Own-cell store through the array's buffer descriptor, then the z = max
// face cell above when this is the top cell (see put_with_zface).
template <int Q, int MODE, class R>
PF_DEV void put_tile(const Mesh<R>& m, const Bufs& bf, unsigned boff, unsigned soff, int i, int j, int k,
                     V<R> s0, V<R> cur, R w) {
    bst(Q == 0 ? bf.n0 : (Q == 1 ? bf.n1 : bf.n2), boff, soff, s0.h);
    if constexpr (MODE == MODE_FP32C) bst(Q == 0 ? bf.nl0 : (Q == 1 ? bf.nl1 : bf.nl2), boff, soff, s0.l);
    const int Lz = Ext<Q>::Lz(m.nz);
    if (k == Lz - 1)
        put<Q, MODE>(m, i, j, Lz, mur_map<MODE>(cur, m.mur[9 + Q], m.mur_lo[9 + Q], s0, cur_q<Q>(m, i, j, Lz, w)));
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
    const int lp = lane + 1 < TK ? lane + 1 : TK - 1;

    for (int e = row * TK + lane; e < m.nmat * MAT_STRIDE; e += TK * TJ) smat[e] = m.mat[e];

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
    const unsigned col = valid ? (unsigned)(j * m.pz + k) : 0u;
    const unsigned boff = col * (unsigned)sizeof(R);
    // non-face cells per component (i-independent part)
    const bool wx = writer && jn && k <= m.nz - 2;
    const bool wy = writer && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2;
    const bool wz = writer && jn && k <= m.nz - 3;

    const Bufs bf = make_bufs(m);
    Col<R> c0 = load_col(m, bf, ib - 1, valid, col, k, w);
    Col<R> nxt = load_col(m, bf, ib, valid, col, k, w);  // prefetched plane i+1
    {
        const int bb = (ib - 1) & 1;
        su[bb][0][row][lane] = c0.x;
        su[bb][1][row][lane] = c0.y;
        su[bb][2][row][lane] = c0.z;
        sid[bb][row][lane] = (uint8_t)c0.id;
    }
    R n1_0 = 0, n2_0 = 0, n3_0 = 0;  // N at plane i
    R s5_m = 0, s6_m = 0;            // T5, T6 at plane i-1

    for (int i = ib - 1; i < ie; ++i) {
        const int ip = i + 1;
        const int b = ip & 1;
        const Col<R> c1 = nxt;
        nxt = load_col(m, bf, ip + 1, valid, col, k, w);  // in flight across this plane
        // previous-level values the update of plane i needs, also in flight
        const unsigned soff = (unsigned)(i < 0 ? 0 : i) * (unsigned)m.plane * (unsigned)sizeof(R);
        const R o0 = bld(bf.n0, boff, soff, R(0));
        const R o1 = bld(bf.n1, boff, soff, R(0));
        const R o2 = bld(bf.n2, boff, soff, R(0));
        R ol0 = 0, ol1 = 0, ol2 = 0, cl0 = 0, cl1 = 0, cl2 = 0;
        if constexpr (COMP) {
            ol0 = bld(bf.nl0, boff, soff, R(0));
            ol1 = bld(bf.nl1, boff, soff, R(0));
            ol2 = bld(bf.nl2, boff, soff, R(0));
            cl0 = bld(bf.cl0, boff, soff, R(0));
            cl1 = bld(bf.cl1, boff, soff, R(0));
            const R c2 = bld(bf.cl2, boff, soff, R(0));
            cl2 = (i == 0 && k == 0) ? R(0) : c2;
        }
        su[b][0][row][lane] = c1.x;
        su[b][1][row][lane] = c1.y;
        su[b][2][row][lane] = c1.z;
        sid[b][row][lane] = (uint8_t)c1.id;
        __syncthreads();

        const R uy1_jm = su[b][1][rm][lane];
        const R uz1_km = from_lane_below(c1.z);
        const R uy0_kp = from_lane_above(c0.y);
        const R ux0_kp = from_lane_above(c0.x);
        const R uz0_jp = su[1 - b][2][rp][lane];
        const R ux0_jp = su[1 - b][0][rp][lane];
        const int id_jp = sid[1 - b][rp][lane];     // (i, j+1, k)
        const int id_kp = sid[1 - b][row][lp];      // (i, j, k+1)
        const int id_jkp = sid[1 - b][rp][lp];      // (i, j+1, k+1)
        const int id1_kp = sid[b][row][lp];         // (i+1, j, k+1)
        const int id1_jp = sid[b][rp][lane];        // (i+1, j+1, k)

        // normal stresses at plane ip (base_solver.py:329-351, 408-418)
        R n1_1, n2_1, n3_1;
        {
            const R* c = smat + MAT_STRIDE * c1.id;
            const R hx = sel(ks, ip >= 1 ? m.sx[ip - 1] : R(1), m.sx0);
            normal3<EXACT>(c, c1.x - c0.x, c1.y - uy1_jm, c1.z - sel(ks, uz1_km, R(0)), hx, hy_n, hz_n,
                           n1_1, n2_1, n3_1);
            const bool nv = ip >= 1 && ip <= m.nx - 2 && jn && kf;
            n1_1 = sel(nv, n1_1, R(0));
            n2_1 = sel(nv, n2_1, R(0));
            n3_1 = sel(nv && ks, n3_1, R(0));
        }
        // shear stresses at plane i (base_solver.py:353-372, 420-433)
        const bool ii = i >= 0 && i <= m.nx - 2;
        const R fxi = sel(ks, ii ? m.fx[i] : R(1), m.fz0);
        R s4 = shear<EXACT>(smat[MAT_STRIDE * (ks ? id_jkp : id_jp) + M_C33], uy0_kp - c0.y, s4_h1,
                            uz0_jp - c0.z, s4_h2);
        R s5 = shear<EXACT>(smat[MAT_STRIDE * (ks ? id1_kp : c1.id) + M_C44], ux0_kp - c0.x, s5_h1,
                            c1.z - c0.z, fxi);
        R s6 = shear<EXACT>(smat[MAT_STRIDE * id1_jp + M_C55], ux0_jp - c0.x, s6_h1, c1.y - c0.y, fxi);
        s4 = sel(ii && i >= 1 && jf && kf, s4, R(0));
        s5 = sel(ii && jn && kf, s5, R(0));
        s6 = sel(ii && jf && kf, s6, R(0));
        sn2[b][row][lane] = n2_1;
        ss4[row][lane] = s4;
        ss6[row][lane] = s6;
        __syncthreads();

        const R n3_0_kp = from_lane_above(n3_0);
        const R s4_km = from_lane_below(s4);
        const R s5_km = from_lane_below(s5);
        if (i >= ib) {
            const R n2_0_jp = sn2[1 - b][rp][lane];
            const R s4_jm = ss4[rm][lane];
            const R s6_jm = ss6[rm][lane];
            const R s = SPONGE ? damp_at(m, i, j < 0 ? 0 : (j < m.ny ? j : m.ny - 1),
                                         k < 0 ? 0 : (k < m.nz ? k : m.nz - 1))
                               : R(0);
            const R gx = sel(ks, i >= 1 && i <= m.nx - 2 ? m.sx[i - 1] : R(1), m.sx0);
            // ux (base_solver.py:441-447, 495-501)
            {
                const R dv = div3<EXACT>(n1_1 - n1_0, sel(ks, i <= m.nx - 2 ? m.fx[i] : R(1), m.fx0),
                                         s6 - s6_jm, hy_n, s5 - sel(ks, s5_km, R(0)), hz_n);
                const V<R> cur = {c0.x, cl0};
                const V<R> s0 = leap_from<MODE, SPONGE>(cur, o0, ol0, smat[MAT_STRIDE * c1.id + M_COEF], dv, s);
                if (wx && i <= m.nx - 3) put_tile<0, MODE>(m, bf, boff, soff, i, j, k, s0, cur, w);
            }
            // uy (base_solver.py:449-455, 503-509)
            {
                const R dv = div3<EXACT>(s6 - s6_m, gx, n2_0_jp - n2_0, uy_g2, s4 - sel(ks, s4_km, R(0)), hz_n);
                const V<R> cur = {c0.y, cl1};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, o1, ol1, smat[MAT_STRIDE * id_jp + M_COEF], dv, s);
                if (i == 0) s0 = {R(0), R(0)};
                if (wy && i <= m.nx - 2) put_tile<1, MODE>(m, bf, boff, soff, i, j, k, s0, cur, w);
            }
            // uz (base_solver.py:457-463, 511-517)
            {
                const R dv = div3<EXACT>(s5 - s5_m, gx, s4 - s4_jm, hy_n, n3_0_kp - n3_0, uz_g3);
                const V<R> cur = {c0.z, cl2};
                V<R> s0 = leap_from<MODE, SPONGE>(cur, o2, ol2, smat[MAT_STRIDE * (ks ? id_kp : c0.id) + M_COEF],
                                                  dv, s);
                if (i == 0) s0 = {R(0), R(0)};
                if (wz && i <= m.nx - 2) put_tile<2, MODE>(m, bf, boff, soff, i, j, k, s0, cur, w);
            }
        }
        c0 = c1;
        n1_0 = n1_1;
        n2_0 = n2_1;
        n3_0 = n3_1;
        s5_m = s5;
        s6_m = s6;
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
