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

template <class R>
PF_DEV Col<R> load_col(const Mesh<R>& m, int ip, bool valid, unsigned col, int k, R w) {
    Col<R> c = {R(0), R(0), R(0), 0};
    if (valid && ip >= 0 && ip < m.nx) {
        const long long base = (long long)ip * m.plane;
        c.x = (m.c0 + base)[col];
        c.y = (m.c1 + base)[col];
        c.z = (ip == 0 && k == 0) ? w : (m.c2 + base)[col];
        c.id = (m.ids + base)[col];
    }
    return c;
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
    const int j = jt * (TJ - 2) - 1 + row;
    const int ib = xc * tm.xchunk;
    const int ie = min(ib + tm.xchunk, m.nx);
    const bool valid = k >= 0 && k < m.nz && j >= 0 && j < m.ny;
    const bool writer = valid && lane >= 1 && lane <= TK - 2 && row >= 1 && row <= TJ - 2;
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;  // uniform over the block
    const R w = (R)m.wave[tt];
    const int rp = row + 1 < TJ ? row + 1 : TJ - 1;
    const int rm = row > 0 ? row - 1 : 0;
    const int lp = lane + 1 < TK ? lane + 1 : TK - 1;

    for (int e = row * TK + lane; e < m.nmat * MAT_STRIDE; e += TK * TJ) smat[e] = m.mat[e];

    // per-column spacing factors (spacings in FP64 mode, inverses otherwise)
    const bool jn = j >= 1 && j <= m.ny - 2;   // staggered y index valid
    const bool jf = j >= 0 && j <= m.ny - 2;   // full y index valid
    const bool kn = k >= 1 && k <= m.nz - 2;
    const bool kf = k >= 0 && k <= m.nz - 2;
    const R hsy = jn ? m.sy[j - 1] : R(1);
    const R hfy = jf ? m.fy[j] : R(1);
    const R hsz = kn ? m.sz[k - 1] : R(1);
    const R hfz = kf ? m.fz[k] : R(1);
    const unsigned col = valid ? (unsigned)(j * m.pz + k) : 0u;

    // plane i state (registers) and plane ib-1 in LDS
    Col<R> c0 = load_col(m, ib - 1, valid, col, k, w);
    Col<R> nxt = load_col(m, ib, valid, col, k, w);  // prefetched plane i+1
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
        nxt = load_col(m, ip + 1, valid, col, k, w);  // in flight across this plane
        // previous-level values the update of plane i needs, also in flight
        const bool out = writer && i >= ib;
        const long long pbase = (long long)(i < 0 ? 0 : i) * m.plane;
        R o0 = 0, o1 = 0, o2 = 0, ol0 = 0, ol1 = 0, ol2 = 0, cl0 = 0, cl1 = 0, cl2 = 0;
        if (out) {
            o0 = (m.n0 + pbase)[col];
            o1 = (m.n1 + pbase)[col];
            o2 = (m.n2 + pbase)[col];
            if constexpr (COMP) {
                ol0 = (m.nl0 + pbase)[col];
                ol1 = (m.nl1 + pbase)[col];
                ol2 = (m.nl2 + pbase)[col];
                cl0 = (m.cl0 + pbase)[col];
                cl1 = (m.cl1 + pbase)[col];
                cl2 = (i == 0 && k == 0) ? R(0) : (m.cl2 + pbase)[col];
            }
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
        R n1_1 = 0, n2_1 = 0, n3_1 = 0;
        if (ip >= 1 && ip <= m.nx - 2 && jn && kf) {
            const R* c = smat + MAT_STRIDE * c1.id;
            const R dux = c1.x - c0.x, duy = c1.y - uy1_jm;
            if (k >= 1) normal3<EXACT>(c, dux, duy, c1.z - uz1_km, m.sx[ip - 1], hsy, hsz, n1_1, n2_1, n3_1);
            else normal3_surface<EXACT>(c, dux, duy, c1.z, m.sx0, m.sy0, m.sz0, n1_1, n2_1, n3_1);
        }
        // shear stresses at plane i (base_solver.py:353-372, 420-433)
        R s4 = 0, s5 = 0, s6 = 0;
        const bool ii = i >= 0 && i <= m.nx - 2;
        if (ii && i >= 1 && jf && kf) {
            if (k >= 1) s4 = shear<EXACT>(smat[MAT_STRIDE * id_jkp + M_C33], uy0_kp - c0.y, hfz, uz0_jp - c0.z, hfy);
            else s4 = shear<EXACT>(smat[MAT_STRIDE * id_jp + M_C33], uy0_kp - c0.y, m.fy0, uz0_jp - c0.z, m.fz0);
        }
        if (ii && jn && kf) {
            if (k >= 1) s5 = shear<EXACT>(smat[MAT_STRIDE * id1_kp + M_C44], ux0_kp - c0.x, hfz, c1.z - c0.z, m.fx[i]);
            else s5 = shear<EXACT>(smat[MAT_STRIDE * c1.id + M_C44], ux0_kp - c0.x, m.fx0, c1.z - c0.z, m.fz0);
        }
        if (ii && jf && kf) {
            const R c = smat[MAT_STRIDE * id1_jp + M_C55];
            if (k >= 1) s6 = shear<EXACT>(c, ux0_jp - c0.x, hfy, c1.y - c0.y, m.fx[i]);
            else s6 = shear<EXACT>(c, ux0_jp - c0.x, m.fx0, c1.y - c0.y, m.fz0);
        }
        sn2[b][row][lane] = n2_1;
        ss4[row][lane] = s4;
        ss6[row][lane] = s6;
        __syncthreads();

        const R n3_0_kp = from_lane_above(n3_0);
        const R s4_km = from_lane_below(s4);
        const R s5_km = from_lane_below(s5);
        if (out) {
            const R n2_0_jp = sn2[1 - b][rp][lane];
            const R s4_jm = ss4[rm][lane];
            const R s6_jm = ss6[rm][lane];
            const R s = SPONGE ? damp_at(m, i, j, k) : R(0);
            // ux (base_solver.py:441-447, 495-501)
            if (i <= m.nx - 3 && jn && k <= m.nz - 2) {
                R dv;
                if (k >= 1) dv = div3<EXACT>(n1_1 - n1_0, m.fx[i], s6 - s6_jm, hsy, s5 - s5_km, hsz);
                else dv = div3<EXACT>(n1_1 - n1_0, m.fx0, s6 - s6_jm, m.sy0, s5 - R(0), m.sz0);
                const R coef = smat[MAT_STRIDE * c1.id + M_COEF];
                const V<R> cur = {c0.x, cl0};
                put_with_zface<0, MODE>(m, i, j, k, leap_from<MODE, SPONGE>(cur, o0, ol0, coef, dv, s), cur, w);
            }
            // uy (base_solver.py:449-455, 503-509)
            if (i <= m.nx - 2 && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2) {
                const V<R> cur = {c0.y, cl1};
                V<R> s0 = {R(0), R(0)};
                if (i > 0) {
                    R dv;
                    if (k >= 1) dv = div3<EXACT>(s6 - s6_m, m.sx[i - 1], n2_0_jp - n2_0, hfy, s4 - s4_km, hsz);
                    else dv = div3<EXACT>(s6 - s6_m, m.sx0, n2_0_jp - n2_0, m.fy0, s4 - R(0), m.sz0);
                    s0 = leap_from<MODE, SPONGE>(cur, o1, ol1, smat[MAT_STRIDE * id_jp + M_COEF], dv, s);
                }
                put_with_zface<1, MODE>(m, i, j, k, s0, cur, w);
            }
            // uz (base_solver.py:457-463, 511-517)
            if (i <= m.nx - 2 && jn && k <= m.nz - 3) {
                const V<R> cur = {c0.z, cl2};
                V<R> s0 = {R(0), R(0)};
                if (i > 0) {
                    R dv, coef;
                    if (k >= 1) {
                        dv = div3<EXACT>(s5 - s5_m, m.sx[i - 1], s4 - s4_jm, hsy, n3_0_kp - n3_0, hfz);
                        coef = smat[MAT_STRIDE * id_kp + M_COEF];
                    } else {
                        dv = div3_uz_surface<EXACT>(s5 - s5_m, m.sx0, s4 - s4_jm, m.sy0, n3_0_kp, n3_0, m.fz0);
                        coef = smat[MAT_STRIDE * c0.id + M_COEF];
                    }
                    s0 = leap_from<MODE, SPONGE>(cur, o2, ol2, coef, dv, s);
                }
                put_with_zface<2, MODE>(m, i, j, k, s0, cur, w);
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
