// Two-step tiled sweep: levels n+1 and n+2 in one pass over the mesh.
//
// The block marches through its chunk of x planes as the one-step tiled
// sweep does (stage A: level n+1 of plane i from levels n and n-1), and a
// second stage B follows one plane behind: level n+2 of plane i-1 from the
// level n+1 values stage A has computed (kept in registers and LDS, never
// re-read from memory) and level n (the node's own value, still in
// registers from stage A).  Per two steps the pass reads levels n and n-1
// and the material indices once and writes levels n+1 and n+2 once: 49 B
// per node in FP32 where two one-step sweeps move 2 x 37.
//
// Tiles overlap by two lanes / rows on each side.  Stage A is exact on all
// but the outermost ring of the tile (lanes 0 and KZ-1, rows 0 and TJ-1 see
// the DPP fill and the clamped LDS rows), stage B on all but the outer two
// rings; those inner cells are the ones the block owns and stores, of both
// levels, so every cell is stored by exactly one block.  Levels n and n-1
// are only read and levels n+1 and n+2 go to two other level buffers, so
// no block reads what another one writes.
//
// Boundary faces.  The z = max face cells of level n+1 are a single Mur map
// of the cell below (which stage A writes with it), so stage B takes them
// from the lane below and covers the top of the columns.  The x and y faces
// are final only after the face kernel, so stage B stores no cell whose
// stencil reads one of them: it stores i <= nx-4, 2 <= j <= ny-4.  The
// one-step tiled march updates the remaining band (k_band) after the
// level n+1 faces, then the level n+2 faces are written.
//
// Stage B sees level n+1 exactly as a one-step sweep loads it: 0 outside
// the mesh and on never-written cells, the source value w(tt+1) on the wall
// line uz(0, j, 0), lo parts rounded to their 16-bit storage.  Every
// formula is the shared per-value code of the one-step sweep fed the same
// operands, so a pass gives the bits of two one-step sweeps.
//
// Barriers: three per plane.  Per plane i (H = LDS parity of plane i+1):
//   A1  N at plane i+1, T at plane i from level n       -> sn2[H], ss4, ss6
//   (a)
//   A2  level n+1 of plane i; stores                     -> sb[H]
//   (b)
//   B1  N at plane i, T at plane i-1 from level n+1      -> sbn2[H], sbs4, sbs6
//       level-n column of plane i+2                      -> su[1-H], sid
//   (c)
//   B2  level n+2 of plane i-1; stores
// Every LDS array is written in one interval and read in the one or two
// after it, before its next write (sid has four slots, plane & 3, so the
// write of plane i+2 in B1 does not touch planes i-1 .. i+1).  A version
// with stage B two planes behind needs only two barriers per plane but
// more registers than a 1024-thread block has (it spilled and ran slower).
#pragma once
#include "pf_tiled.hpp"

namespace pf {

constexpr int XT2 = 256;  // x-table entries of the two-step sweep: planes per chunk + 6

template <class R, int TJ, int KZ>
struct FusedLds {
    R su[2][3][TJ][KZ];    // level n planes, by plane parity
    R sn2[2][TJ][KZ];      // stage A: T2 by plane parity
    R ss4[TJ][KZ];         // stage A: T4, T6 of the current plane
    R ss6[TJ][KZ];
    R sb[2][3][TJ][KZ];    // level n+1 planes, by plane parity
    R sbn2[2][TJ][KZ];     // stage B: T2 by plane parity
    R sbs4[TJ][KZ];        // stage B: T4, T6 of its current plane
    R sbs6[TJ][KZ];
    uint16_t sid[4][TJ][KZ];  // material row offsets, by plane & 3
    R smat[PF_TILED_MAX_MATERIALS * MAT_STRIDE];
    XPlane<R> sxp[XT2];    // planes ib-2 .. ie+3
};

// Level n+1 of a node as a one-step sweep loads it: hi as stored, lo
// rounded to its 16-bit storage, 0 off the mesh.
template <int MODE, class R>
PF_DEV V<R> as_stored(V<R> v, bool in) {
    V<R> r;
    r.h = in ? v.h : R(0);
    if constexpr (MODE == MODE_FP32C) r.l = in ? lo_decode(v.h, lo_encode(v.h, v.l)) : R(0);
    else r.l = R(0);
    return r;
}

template <class R, bool COMP>
PF_DEV Rsrc level_rsrc(const Mesh<R>& m, R* base) {
    const long long n = (long long)m.nx * m.plane;
    const long long lvl = n * 3 * (long long)sizeof(R) + (COMP ? n * 3 * (long long)sizeof(lo_t) : 0);
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)(unsigned)lvl, PF_RSRC_FLAGS);
}

// The march of one block over its tile and chunk.  m is the view of the
// first step (level n -> n+1, written to m.n*), out2 the level buffer of
// n+2 (same layout).  FULL: every lane of the tile lies at 1 <= k <= nz-5
// and 1 <= j <= ny-5 and the chunk at 3 <= ib, ie <= nx-4, so no lane meets
// the free surface, a range limit, a face or the band, and the range
// selects of the general code are the identity (same bits).
template <class R, int MODE, bool SPONGE, int TJ, int KZ, bool FULL>
PF_DEV void fused_march(const Mesh<R>& m, R* out2, FusedLds<R, TJ, KZ>& L, int kl, int row, int k, int j, int ib,
                        int ie, R w, R w1) {
    constexpr int EXACT = exact_of(MODE);
    constexpr bool COMP = MODE == MODE_FP32C;
    constexpr int RB = MAT_STRIDE * (int)sizeof(R);
    const int rp = row + 1 < TJ ? row + 1 : TJ - 1;
    const int rm = row > 0 ? row - 1 : 0;

    // column constants (as in the one-step sweep)
    const bool kval = FULL || (k >= 0 && k < m.nz);
    const bool jval = FULL || (j >= 0 && j < m.ny);
    const bool valid = kval && jval;
    const bool ks = FULL || k >= 1;
    const bool jn = FULL || (j >= 1 && j <= m.ny - 2);
    const bool jf = FULL || (j >= 0 && j <= m.ny - 2);
    const bool kn = FULL || (k >= 1 && k <= m.nz - 2);
    const bool kf = FULL || (k >= 0 && k <= m.nz - 2);
    const bool own = valid && kl >= 2 && kl <= KZ - 3 && row >= 2 && row <= TJ - 3;
    const R hsy = jn ? m.sy[j - 1] : R(1);
    const R hfy = jf ? m.fy[j] : R(1);
    const R hsz = kn ? m.sz[k - 1] : R(1);
    const R hfz = kf ? m.fz[k] : R(1);
    const R hy_n = sel(ks, hsy, m.sy0);
    const R hz_n = sel(ks, hsz, m.sz0);
    const R s4_h1 = sel(ks, hfz, m.fy0), s4_h2 = sel(ks, hfy, m.fz0);
    const R s5_h1 = sel(ks, hfz, m.fx0);
    const R s6_h1 = sel(ks, hfy, m.fx0);
    const R uy_g2 = sel(ks, hfy, m.fy0);
    const R uz_g3 = sel(ks || m.surf, hfz, R(1));  // as in the one-step sweep
    R dpyz = R(0), dpz_k = R(0);
    if constexpr (SPONGE) {
        dpyz = m.dpy[j < 0 ? 0 : (j < m.ny ? j : m.ny - 1)];
        dpz_k = m.dpz[k < 0 ? 0 : (k < m.nz ? k : m.nz - 1)];
    }
    const unsigned col = valid ? (unsigned)(j * m.pz + k) : 0u;
    const unsigned boff = col * (unsigned)sizeof(R);
    const unsigned loff = col * (unsigned)sizeof(lo_t);
    const unsigned plane = (unsigned)m.plane;
    // level n+1 stores: owned non-face cells, z = max face above the top
    const bool wx = own && jn && k <= m.nz - 2;
    const bool wy = own && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2;
    const bool wz = own && jn && k <= m.nz - 3;
    const bool top_xy = !FULL && k == m.nz - 2;
    const bool top_z = !FULL && k == m.nz - 3;
    // level n+2 is stored off the x and y bands
    const bool wb = FULL || (j >= 2 && j <= m.ny - 4);

    const Bufs bf = make_bufs<R, COMP>(m);  // cur = n, old = n-1, nxt = n+1
    Bufs bf2 = bf;
    bf2.nxt = level_rsrc<R, COMP>(m, out2);  // n+2

    // stage A's values are used on all but the outer ring of the tile
    const bool useo = valid && kl >= 1 && kl <= KZ - 2 && row >= 1 && row <= TJ - 2;
    const unsigned ooff = useo || !bf.far ? boff : bf.far;
    const unsigned oloff = useo || !bf.far ? loff : bf.far;
    const int pmax = ie + 1 < m.nx ? ie + 1 : m.nx - 1;  // last level-n plane read
    const int omax = ie < m.nx ? ie : m.nx - 1;          // last plane whose node values are read

    // prologue: level-n planes ib-2, ib-1 into registers and LDS, the loads
    // of planes ib, ib+1 and the node values of plane ib-1 in flight
    Col<R> c0 = settle(load_raw<R>(bf, ib >= 2 ? ib - 2 : 0, plane, boff, col), valid && ib >= 2,
                       !FULL && ib == 2 && k == 0, w);
    Col<R> c1 = settle(load_raw<R>(bf, ib >= 1 ? ib - 1 : 0, plane, boff, col), valid && ib >= 1,
                       !FULL && ib == 1 && k == 0, w);
    c0.id *= RB;
    c1.id *= RB;
    Raw<R> ra = load_raw<R>(bf, ib, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Raw<R> rb = load_raw<R>(bf, ib + 1 <= pmax ? ib + 1 : pmax, plane, boff, col);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> ob = load_old<R, COMP>(bf, ib >= 1 ? ib - 1 : 0, plane, boff, loff, ooff, oloff);
    __builtin_amdgcn_sched_barrier(0);
    Old<R> oa = {};  // the first plane (ib-2) is warm-up only
    L.su[1][0][row][kl] = c0.x;
    L.su[1][1][row][kl] = c0.y;
    L.su[1][2][row][kl] = c0.z;
    L.sid[(ib - 2) & 3][row][kl] = (uint16_t)c0.id;
    L.su[0][0][row][kl] = c1.x;
    L.su[0][1][row][kl] = c1.y;
    L.su[0][2][row][kl] = c1.z;
    L.sid[(ib - 1) & 3][row][kl] = (uint16_t)c1.id;
    L.sid[ib & 3][row][kl] = 0;  // read (as plane ib-3) before any plane is there
    L.sid[(ib + 1) & 3][row][kl] = 0;
    __syncthreads();

    // stage A: N at plane i, T5/T6 at plane i-1
    R n1_0 = 0, n2_0 = 0, n3_0 = 0, s5_m = 0, s6_m = 0;
    // stage B: level n+1 of plane i-1 and its material row, N (level n+1)
    // at plane i-1, T5/T6 at plane i-2, level n of plane i-1
    V<R> b0x = {R(0), R(0)}, b0y = {R(0), R(0)}, b0z = {R(0), R(0)};
    int b0id = 0;
    R bn1_0 = 0, bn2_0 = 0, bn3_0 = 0, bs5_m = 0, bs6_m = 0;
    V<R> px = {R(0), R(0)}, py = {R(0), R(0)}, pz = {R(0), R(0)};
    const R* smat = L.smat;

    // One plane i.  rin holds the raw level-n column of plane i+2 and
    // receives plane i+4; oin holds the node values of plane i and receives
    // plane i+2.  c0, c1: level n of planes i, i+1 (also in su[1-H], su[H]).
    auto plane_step = [&](int i, auto Hc, Raw<R>& rin, Old<R>& oin) {
        constexpr int H = decltype(Hc)::value;
        const int ip = i + 1;
        const int q0 = i & 3, q1 = ip & 3, qm = (i - 1) & 3, q2 = (i + 2) & 3;

        // ---- A1: N at plane i+1, T at plane i (level n) ----
        const XPlane<R> xq0 = L.sxp[i - ib + 2];
        const XPlane<R> xq1 = L.sxp[ip - ib + 2];
        const R uy1_jm = L.su[H][1][rm][kl];
        const R uz1_km = below0(c1.z);
        const R uy0_kp = above0(c0.y);
        const R ux0_kp = above0(c0.x);
        const R uz0_jp = L.su[1 - H][2][rp][kl];
        const R ux0_jp = L.su[1 - H][0][rp][kl];
        const int id_jp = L.sid[q0][rp][kl];
        const int id1_jp = L.sid[q1][rp][kl];
        const int id_kp = dpp_above0(c0.id);
        const int id_jkp = dpp_above0(id_jp);
        const int id1_kp = dpp_above0(c1.id);
        R n1_1, n2_1, n3_1;
        normal3<EXACT>(mrow(smat, c1.id), c1.x - c0.x, c1.y - uy1_jm, c1.z - sel(ks, uz1_km, R(0)),
                       sel(ks, xq1.hs, m.sx0), hy_n, hz_n, n1_1, n2_1, n3_1);
        if constexpr (!FULL) {
            const bool nv = ip >= 1 && ip <= m.nx - 2 && jn && kf;
            n1_1 = sel(nv, n1_1, R(0));
            n2_1 = sel(nv, n2_1, R(0));
            n3_1 = sel(nv && ks, n3_1, R(0));
        }
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
        __syncthreads();  // (a)

        // ---- A2: level n+1 of plane i ----
        const R n3_0_kp = above0(n3_0);
        const R s4_km = below0(s4);
        const R s5_km = below0(s5);
        const R n2_0_jp = L.sn2[1 - H][rp][kl];
        const R s4_jm = L.ss4[rm][kl];
        const R s6_jm = L.ss6[rm][kl];
        const R sa = SPONGE ? (xq0.dp + dpyz) + dpz_k : R(0);
        const R gxa = sel(ks, xq0.hs, m.sx0);
        const bool sta = i >= ib && i < ie;  // plane this block stores
        const unsigned soff = uni((unsigned)i * plane * (unsigned)sizeof(R));
        const unsigned slo = uni((unsigned)i * plane * (unsigned)sizeof(lo_t));
        // lo parts decoded against their stored hi (c0 is the stored value
        // except at the wall source line, whose lo is 0)
        const R cl0 = COMP ? lo_decode(c0.x, oin.cl0) : R(0);
        const R cl1 = COMP ? lo_decode(c0.y, oin.cl1) : R(0);
        const R cl2 = (!COMP || (!FULL && i == 0 && k == 0)) ? R(0) : lo_decode(c0.z, oin.cl2);
        const R ol0 = COMP ? lo_decode(oin.o0, oin.ol0) : R(0);
        const R ol1 = COMP ? lo_decode(oin.o1, oin.ol1) : R(0);
        const R ol2 = COMP ? lo_decode(oin.o2, oin.ol2) : R(0);
        const bool wall = !FULL && i == 0;
        const V<R> curx = {c0.x, cl0}, cury = {c0.y, cl1}, curz = {c0.z, cl2};
        const V<R> abx = {ux0_kp, COMP ? above0(cl0) : R(0)};
        const V<R> aby = {uy0_kp, COMP ? above0(cl1) : R(0)};
        const V<R> abz = {above0(c0.z), COMP ? above0(cl2) : R(0)};
        V<R> ax, ay, az;
        {
            const R dv = div3<EXACT>(n1_1 - n1_0, sel(ks, xq0.hf, m.fx0), s6 - s6_jm, hy_n, s5 - sel(ks, s5_km, R(0)),
                                     hz_n);
            ax = leap_from<MODE, SPONGE>(curx, oin.o0, ol0, mget(smat, c1.id, M_COEF), dv, sa);
            if (sta && wx && (FULL || i <= m.nx - 3))
                put_tile<0, MODE>(m, bf, boff, loff, soff, slo, top_xy, ax, curx, abx);
        }
        {
            const R dv = div3<EXACT>(s6 - s6_m, gxa, n2_0_jp - n2_0, uy_g2, s4 - sel(ks, s4_km, R(0)), hz_n);
            ay = leap_from<MODE, SPONGE>(cury, oin.o1, ol1, mget(smat, id_jp, M_COEF), dv, sa);
            if (wall) ay = {R(0), R(0)};
            if (sta && wy && (FULL || i <= m.nx - 2))
                put_tile<1, MODE>(m, bf, boff, loff, soff, slo, top_xy, ay, cury, aby);
        }
        {
            const R dv = div3<EXACT>(s5 - s5_m, gxa, s4 - s4_jm, hy_n, n3_0_kp - n3_0, uz_g3);
            az = leap_from<MODE, SPONGE>(curz, oin.o2, ol2, mget(smat, ks ? id_kp : c0.id, M_COEF), dv, sa);
            if (wall) az = {R(0), R(0)};
            if (sta && wz && (FULL || i <= m.nx - 2))
                put_tile<2, MODE>(m, bf, boff, loff, soff, slo, top_z, az, curz, abz);
        }
        oin = load_old<R, COMP>(bf, i + 2 <= omax ? i + 2 : omax, plane, boff, loff, ooff, oloff);
        // level n+1 of plane i as stage B reads it: as stored; on the
        // z = max faces the face value the lane below computes (a Mur map of
        // its unrounded stage-I value, as put_tile stores it), 0 on the
        // never-written uz(i, j, nz-1); the wall line takes w(tt+1)
        const bool pin = valid && (FULL || (i >= 0 && i < m.nx));
        V<R> b1x, b1y, b1z;
        if constexpr (!FULL) {
            const V<R> fx = mur_map<MODE>(curx, m.mur[9], m.mur_lo[9], ax, abx);
            const V<R> fy = mur_map<MODE>(cury, m.mur[10], m.mur_lo[10], ay, aby);
            const V<R> fz = mur_map<MODE>(curz, m.mur[11], m.mur_lo[11], az, abz);
            // the lane shifts run on every lane, outside any condition
            const V<R> gx = {below0(fx.h), COMP ? below0(fx.l) : R(0)};
            const V<R> gy = {below0(fy.h), COMP ? below0(fy.l) : R(0)};
            const V<R> gz = {below0(fz.h), COMP ? below0(fz.l) : R(0)};
            const bool zx = k == m.nz - 1, zz = k == m.nz - 2;
            b1x = as_stored<MODE>(V<R>{zx ? gx.h : ax.h, zx ? gx.l : ax.l}, pin);
            b1y = as_stored<MODE>(V<R>{zx ? gy.h : ay.h, zx ? gy.l : ay.l}, pin);
            b1z = as_stored<MODE>(V<R>{zz ? gz.h : az.h, zz ? gz.l : az.l}, pin && !zx);
            if (i == 0 && k == 0) b1z = {valid ? w1 : R(0), R(0)};
        } else {
            b1x = as_stored<MODE>(ax, pin);
            b1y = as_stored<MODE>(ay, pin);
            b1z = as_stored<MODE>(az, pin);
        }
        L.sb[H][0][row][kl] = b1x.h;
        L.sb[H][1][row][kl] = b1y.h;
        L.sb[H][2][row][kl] = b1z.h;
        __syncthreads();  // (b)

        // ---- B1: N at plane i, T at plane i-1 (level n+1) ----
        const int p = i - 1;
        const int b1id = c0.id;
        const XPlane<R> xb0 = L.sxp[i - ib + 1 > 0 ? i - ib + 1 : 0];  // plane p
        const R vy1_jm = L.sb[H][1][rm][kl];
        const R vz1_km = below0(b1z.h);
        const R vy0_kp = above0(b0y.h);
        const R vx0_kp = above0(b0x.h);
        const R vz0_jp = L.sb[1 - H][2][rp][kl];
        const R vx0_jp = L.sb[1 - H][0][rp][kl];
        const int bid_jp = L.sid[qm][rp][kl];
        const int bid1_jp = L.sid[q0][rp][kl];
        const int bid_kp = dpp_above0(b0id);
        const int bid_jkp = dpp_above0(bid_jp);
        const int bid1_kp = dpp_above0(b1id);
        R bn1_1, bn2_1, bn3_1;
        normal3<EXACT>(mrow(smat, b1id), b1x.h - b0x.h, b1y.h - vy1_jm, b1z.h - sel(ks, vz1_km, R(0)),
                       sel(ks, xq0.hs, m.sx0), hy_n, hz_n, bn1_1, bn2_1, bn3_1);
        if constexpr (!FULL) {
            const bool nv = i >= 1 && i <= m.nx - 2 && jn && kf;
            bn1_1 = sel(nv, bn1_1, R(0));
            bn2_1 = sel(nv, bn2_1, R(0));
            bn3_1 = sel(nv && ks, bn3_1, R(0));
        }
        const R fxb = sel(ks, xb0.hf, m.fz0);
        R t4 = shear<EXACT>(mget(smat, ks ? bid_jkp : bid_jp, M_C33), vy0_kp - b0y.h, s4_h1, vz0_jp - b0z.h, s4_h2);
        R t5 = shear<EXACT>(mget(smat, ks ? bid1_kp : b1id, M_C44), vx0_kp - b0x.h, s5_h1, b1z.h - b0z.h, fxb);
        R t6 = shear<EXACT>(mget(smat, bid1_jp, M_C55), vx0_jp - b0x.h, s6_h1, b1y.h - b0y.h, fxb);
        if constexpr (!FULL) {
            const bool ii = p >= 0 && p <= m.nx - 2;
            t4 = sel(ii && p >= 1 && jf && kf, t4, R(0));
            t5 = sel(ii && jn && kf, t5, R(0));
            t6 = sel(ii && jf && kf, t6, R(0));
        }
        L.sbn2[H][row][kl] = bn2_1;
        L.sbs4[row][kl] = t4;
        L.sbs6[row][kl] = t6;
        // level n of plane i+2 into LDS (its loads were issued two planes ago)
        Col<R> c2 = settle(rin, valid && (FULL || i + 2 < m.nx), !FULL && i + 2 == 0 && k == 0, w);
        c2.id *= RB;
        L.su[1 - H][0][row][kl] = c2.x;
        L.su[1 - H][1][row][kl] = c2.y;
        L.su[1 - H][2][row][kl] = c2.z;
        L.sid[q2][row][kl] = (uint16_t)c2.id;
        __syncthreads();  // (c)
        // issued after the barrier, so the registers of rin are free again
        rin = load_raw<R>(bf, i + 4 <= pmax ? i + 4 : pmax, plane, boff, col);

        // ---- B2: level n+2 of plane i-1 ----
        {
            const R bn3_0_kp = above0(bn3_0);
            const R t4_km = below0(t4);
            const R t5_km = below0(t5);
            const R bn2_0_jp = L.sbn2[1 - H][rp][kl];
            const R t4_jm = L.sbs4[rm][kl];
            const R t6_jm = L.sbs6[rm][kl];
            const R sb_ = SPONGE ? (xb0.dp + dpyz) + dpz_k : R(0);
            const R gxb = sel(ks, xb0.hs, m.sx0);
            const bool stb = p >= ib && p < ie && wb && (FULL || p <= m.nx - 4);
            const V<R> bax = {vx0_kp, COMP ? above0(b0x.l) : R(0)};
            const V<R> bay = {vy0_kp, COMP ? above0(b0y.l) : R(0)};
            const V<R> baz = {above0(b0z.h), COMP ? above0(b0z.l) : R(0)};
            const unsigned soffb = uni((unsigned)p * plane * (unsigned)sizeof(R));
            const unsigned slob = uni((unsigned)p * plane * (unsigned)sizeof(lo_t));
            const bool wallb = !FULL && p == 0;
            {
                const R dv = div3<EXACT>(bn1_1 - bn1_0, sel(ks, xb0.hf, m.fx0), t6 - t6_jm, hy_n,
                                         t5 - sel(ks, t5_km, R(0)), hz_n);
                const V<R> r = leap_from<MODE, SPONGE>(b0x, px.h, px.l, mget(smat, b1id, M_COEF), dv, sb_);
                if (stb && wx) put_tile<0, MODE>(m, bf2, boff, loff, soffb, slob, top_xy, r, b0x, bax);
            }
            {
                const R dv = div3<EXACT>(t6 - bs6_m, gxb, bn2_0_jp - bn2_0, uy_g2, t4 - sel(ks, t4_km, R(0)), hz_n);
                V<R> r = leap_from<MODE, SPONGE>(b0y, py.h, py.l, mget(smat, bid_jp, M_COEF), dv, sb_);
                if (wallb) r = {R(0), R(0)};
                if (stb && wy) put_tile<1, MODE>(m, bf2, boff, loff, soffb, slob, top_xy, r, b0y, bay);
            }
            {
                const R dv = div3<EXACT>(t5 - bs5_m, gxb, t4 - t4_jm, hy_n, bn3_0_kp - bn3_0, uz_g3);
                V<R> r = leap_from<MODE, SPONGE>(b0z, pz.h, pz.l, mget(smat, ks ? bid_kp : b0id, M_COEF), dv, sb_);
                if (wallb) r = {R(0), R(0)};
                if (stb && wz) put_tile<2, MODE>(m, bf2, boff, loff, soffb, slob, top_z, r, b0z, baz);
            }
        }

        // advance one plane
        b0x = b1x;
        b0y = b1y;
        b0z = b1z;
        b0id = b1id;
        px = curx;
        py = cury;
        pz = curz;
        c0 = c1;
        c1 = c2;
        n1_0 = n1_1;
        n2_0 = n2_1;
        n3_0 = n3_1;
        s5_m = s5;
        s6_m = s6;
        bn1_0 = bn1_1;
        bn2_0 = bn2_1;
        bn3_0 = bn3_1;
        bs5_m = t5;
        bs6_m = t6;
    };

    // planes ib-2 .. ie (one more when the count is odd: it loads and
    // computes, stores nothing)
    for (int i = ib - 2; i <= ie; i += 2) {
        plane_step(i, std::integral_constant<int, 0>(), ra, oa);
        plane_step(i + 1, std::integral_constant<int, 1>(), rb, ob);
    }
}

// Tiles of the two-step sweep: stride KZ-4 lanes by TJ-4 rows.  x chunks
// as tile_map chooses them; a block marches three planes beyond its chunk
// (two warm-up planes, one trailing plane of stage B, rounded to a pair).
inline TileMap fused_map(const pf_ctx* c, int tj, int kz, int xchunk, long long slots = 0) {
    TileMap tm;
    tm.nkt = (c->nz + (kz - 4) - 1) / (kz - 4);
    tm.njt = (c->ny + (tj - 4) - 1) / (tj - 4);
    if (xchunk <= 0 && slots > 0) {
        xchunk = round_xchunk((long long)tm.nkt * tm.njt, c->nx, 3, 32, XT2 - 6, slots);
    } else if (xchunk <= 0) {
        // aim for ~1024 blocks (4 per CU) with chunks of at least 32 planes
        const long long per = (long long)tm.nkt * tm.njt;
        long long want = (1024 + per - 1) / per;
        if (want < 1) want = 1;
        xchunk = (int)((c->nx + want - 1) / want);
        if (xchunk < 32) xchunk = 32;
    }
    if (xchunk > XT2 - 6) xchunk = XT2 - 6;
    if (xchunk > c->nx) xchunk = c->nx;
    tm.xchunk = xchunk;
    tm.nxc = (c->nx + xchunk - 1) / xchunk;
    return tm;
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
PF_DEV void fused_tile(const Mesh<R>& m, R* out2, const TileMap& tm, int t, FusedLds<R, TJ, KZ>& L) {
    const int kl = threadIdx.x % KZ;
    const int row = threadIdx.y * (64 / KZ) + threadIdx.x / KZ;
    const int kt = t % tm.nkt;
    const int r_ = t / tm.nkt;
    const int jt = r_ % tm.njt;
    const int xc = r_ / tm.njt;
    const int k = kt * (KZ - 4) - 2 + kl;
    const int j = jt * (TJ - 4) - 2 + row;
    const int ib = xc * tm.xchunk;
    const int ie = min(ib + tm.xchunk, m.nx);
    const long long tt = step_index(m);
    if (tt + 1 >= m.nsteps) return;  // uniform over the grid
    const R w = (R)m.wave[tt];
    const R w1 = (R)m.wave[tt + 1];
    const int tid = threadIdx.y * 64 + threadIdx.x;

    for (int e = tid; e < m.nmat * MAT_STRIDE; e += TJ * KZ) L.smat[e] = m.mat[e];
    for (int e = tid; e < ie - ib + 6; e += TJ * KZ) {
        const int p = ib - 2 + e;
        XPlane<R> x;
        x.hs = (p >= 1 && p <= m.nx - 2) ? m.sx[p - 1] : R(1);
        x.hf = (p >= 0 && p <= m.nx - 2) ? m.fx[p] : R(1);
        x.dp = (SPONGE && p >= 0 && p < m.nx) ? m.dpx[p] : R(0);
        x.pad = R(0);
        L.sxp[e] = x;
    }
    const int k_lo = kt * (KZ - 4) - 2, k_hi = k_lo + KZ - 1;
    const int j_lo = jt * (TJ - 4) - 2, j_hi = j_lo + TJ - 1;
    const bool full = k_lo >= 1 && k_hi <= m.nz - 5 && j_lo >= 1 && j_hi <= m.ny - 5 && ib >= 3 && ie <= m.nx - 4;
    if (full)
        fused_march<R, MODE, SPONGE, TJ, KZ, true>(m, out2, L, kl, row, k, j, ib, ie, w, w1);
    else
        fused_march<R, MODE, SPONGE, TJ, KZ, false>(m, out2, L, kl, row, k, j, ib, ie, w, w1);
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
__global__ __launch_bounds__(TJ * KZ, 1) void k_sweep_fused(Mesh<R> m, R* out2, TileMap tm) {
    __shared__ FusedLds<R, TJ, KZ> L;
    fused_tile<R, MODE, SPONGE, TJ, KZ>(m, out2, tm, xcd_remap((int)blockIdx.x, (int)gridDim.x), L);
}

// ----------------------------------------------------------- band ----
// Cells the two-step pass leaves to the one-step sweep:
//   J: j in {1, ny-3, ny-2}, i <= nx-2, k <= nz-2
//   I: i in {nx-3, nx-2}, 2 <= j <= ny-4, k <= nz-2
// (no component has a non-face cell at i = nx-1, j = 0, j = ny-1 or
// k = nz-1; the z = max faces above k = nz-2 / nz-3 come with their cell).
// The one-step tiled march sweeps them:
//   J tiles, 4 rows x 64 lanes: two tile rows, rows -1..2 (writer rows 0
//            and 1; row 0 has no non-face cell) and rows ny-4..ny-1
//            (writer rows ny-3, ny-2), over all planes in chunks of 8;
//   I tiles, 8 rows x 32 lanes: every tile row (stride 6, writer rows
//            6 jt .. 6 jt + 5), over the one chunk of the planes nx-3,
//            nx-2 (a chunk of two planes reads four, so taller tiles,
//            which reread fewer halo rows, pay more than longer rows).
// So every tile stores band cells only; the J cells of planes nx-3, nx-2
// are stored by a J and an I tile, from the same inputs with the same
// code, so with the same bits.  Needs nx, ny, nz >= 8.  Each part maps its
// blocks through xcd_remap, so the tiles one XCD runs are neighbours that
// share halo planes and lanes in its L2 (9 % faster at 513x513x257 nodes
// than blocks dealt round-robin; longer J chunks, 16 or 32 planes, and
// six blocks per CU measured slower).
struct BandMap {
    TileMap j, i;
};

inline BandMap band_map(const pf_ctx* c) {
    BandMap b;
    b.j = tile_map(c, 4, 64, 8);
    b.j.njt = 2;
    b.i = tile_map(c, 8, 32, 2);
    b.i.nxc = 1;
    return b;
}

inline long long band_blocks(const BandMap& b) {
    return (long long)b.j.nkt * b.j.njt * b.j.nxc + (long long)b.i.nkt * b.i.njt * b.i.nxc;
}

template <class R, int MODE, bool SPONGE>
__global__ __launch_bounds__(256, (MODE == MODE_FP32 ? 4 : 2)) void k_band(Mesh<R> m, TileMap tj, TileMap ti) {
    __shared__ union {
        TiledLds<R, 4, 64> j;
        TiledLds<R, 8, 32> i;
    } L;
    const int nj = tj.nkt * tj.njt * tj.nxc;
    const int b = (int)blockIdx.x;
    if (b < nj)
        tiled_tile<R, MODE, SPONGE, 4, 64>(m, tj, xcd_remap(b, nj), L.j, m.ny - 3, 0);
    else
        tiled_tile<R, MODE, SPONGE, 8, 32>(m, ti, xcd_remap(b - nj, (int)gridDim.x - nj), L.i, 6, m.nx - 3);
}

}  // namespace pf
