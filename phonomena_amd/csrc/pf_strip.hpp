// Barrier-free strip sweep: each wavefront owns NR rows (y) of a 64-lane
// z tile and marches x; its rows' values live in registers.
//
// A wavefront is 64 lanes along z (lanes 0 and 63 are halo lanes that
// compute the stresses their neighbours need and write nothing).  It holds
// NR output rows j0..j0+NR-1 plus one halo row on each side in registers, so
// y neighbours are register reads and z neighbours are DPP lane shifts
// (v_mov_b32_dpp wave_shr:1 / wave_shl:1).  No LDS exchange and no barrier
// is needed after the one-off staging of the material table, so the waves
// of a block stream independently and memory latency is covered by the
// other waves on the SIMD.  The next plane of every row is loaded one
// iteration ahead; the previous level's values of the output rows are
// loaded at the top of the iteration that consumes them.
//
// Per plane, with rows indexed rr = 0..NR+1 (j = j0 - 1 + rr):
//   N = (T1,T2,T3) at plane i+1 for rr = 1..NR+1 (row NR+1 only feeds T2),
//   T4, T6 at plane i for rr = 0..NR, T5 for rr = 1..NR,
//   the stage-I update of plane i for the output rows rr = 1..NR.
// Same per-value formulas and operands as the naive and tiled sweeps, so
// the results are bit-identical.
#pragma once
#include "pf_tiled.hpp"

namespace pf {

constexpr int KOUT = 62;  // output columns per 64-lane tile row

struct StripMap {
    int nkt, nsg, nxc, xchunk;  // z tiles, strip groups (of W strips), x chunks, planes per chunk
};

template <class R, int MODE, bool SPONGE, int NR, int W>
__global__ __launch_bounds__(64 * W, (MODE == MODE_FP32 ? 4 : 2)) void k_sweep_strip(Mesh<R> m, StripMap sm) {
    constexpr int EXACT = exact_of(MODE);
    constexpr bool COMP = MODE == MODE_FP32C;
    constexpr int NRR = NR + 2;
    __shared__ R smat[PF_TILED_MAX_MATERIALS * MAT_STRIDE];
    for (int e = threadIdx.x; e < m.nmat * MAT_STRIDE; e += 64 * W) smat[e] = m.mat[e];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int kt = t % sm.nkt;
    const int r_ = t / sm.nkt;
    const int sg = r_ % sm.nsg;
    const int xc = r_ / sm.nsg;
    const int j0 = (sg * W + wv) * NR;
    const int k = kt * KOUT - 1 + lane;
    const int ib = xc * sm.xchunk;
    const int ie = min(ib + sm.xchunk, m.nx);
    const long long tt = step_index(m);
    if (tt >= m.nsteps) return;  // uniform over the grid
    if (j0 >= m.ny) return;      // whole strip outside the mesh (wave-uniform)
    const R w = (R)m.wave[tt];
    const bool kval = k >= 0 && k < m.nz;
    const bool klane = lane >= 1 && lane <= 62;
    const bool kn = k >= 1 && k <= m.nz - 2;
    const bool kf = k >= 0 && k <= m.nz - 2;
    const R hsz = kn ? m.sz[k - 1] : R(1);
    const R hfz = kf ? m.fz[k] : R(1);

    // row-uniform data (scalar per wavefront)
    bool rv[NRR];
    bool rjn[NRR], rjf[NRR];
    R hsy[NRR], hfy[NRR];
    unsigned col[NRR];
#pragma unroll
    for (int rr = 0; rr < NRR; ++rr) {
        const int j = j0 - 1 + rr;
        rv[rr] = kval && j >= 0 && j < m.ny;
        rjn[rr] = j >= 1 && j <= m.ny - 2;
        rjf[rr] = j >= 0 && j <= m.ny - 2;
        hsy[rr] = rjn[rr] ? m.sy[j - 1] : R(1);
        hfy[rr] = rjf[rr] ? m.fy[j] : R(1);
        col[rr] = rv[rr] ? (unsigned)(j * m.pz + k) : 0u;
    }

    R x0[NRR], y0[NRR], z0[NRR];
    int d0[NRR];
    R px[NRR], py[NRR], pz[NRR];
    int pd[NRR];
    auto load_plane = [&](int ip, R* ox, R* oy, R* oz, int* od) {
        const bool in = ip >= 0 && ip < m.nx;
        const long long base = (long long)(in ? ip : 0) * m.plane;
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) {
            R a = 0, b = 0, c = 0;
            int d = 0;
            if (in && rv[rr]) {
                a = (m.c0 + base)[col[rr]];
                b = (m.c1 + base)[col[rr]];
                c = (ip == 0 && k == 0) ? w : (m.c2 + base)[col[rr]];
                d = (m.ids + base)[col[rr]];
            }
            ox[rr] = a;
            oy[rr] = b;
            oz[rr] = c;
            od[rr] = d;
        }
    };
    load_plane(ib - 1, x0, y0, z0, d0);
    load_plane(ib, px, py, pz, pd);

    R n1_0[NRR], n2_0[NRR], n3_0[NRR];  // N at plane i (rows 1..NR+1)
    R s5m[NRR], s6m[NRR];               // T5, T6 at plane i-1 (rows 1..NR)
#pragma unroll
    for (int rr = 0; rr < NRR; ++rr) n1_0[rr] = n2_0[rr] = n3_0[rr] = s5m[rr] = s6m[rr] = R(0);

    for (int i = ib - 1; i < ie; ++i) {
        const int ip = i + 1;
        R x1[NRR], y1[NRR], z1[NRR];
        int d1[NRR];
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) {
            x1[rr] = px[rr];
            y1[rr] = py[rr];
            z1[rr] = pz[rr];
            d1[rr] = pd[rr];
        }
        load_plane(ip + 1, px, py, pz, pd);  // in flight across this plane

        const bool outp = i >= ib;
        const long long pbase = (long long)(i < 0 ? 0 : i) * m.plane;
        R o0[NRR], o1[NRR], o2[NRR], ol0[NRR], ol1[NRR], ol2[NRR], cl0[NRR], cl1[NRR], cl2[NRR];
#pragma unroll
        for (int rr = 1; rr <= NR; ++rr) {
            o0[rr] = o1[rr] = o2[rr] = ol0[rr] = ol1[rr] = ol2[rr] = cl0[rr] = cl1[rr] = cl2[rr] = R(0);
            if (outp && rv[rr] && klane) {
                o0[rr] = (m.o0 + pbase)[col[rr]];
                o1[rr] = (m.o1 + pbase)[col[rr]];
                o2[rr] = (m.o2 + pbase)[col[rr]];
                if constexpr (COMP) {
                    ol0[rr] = lo_decode(o0[rr], (m.ol0 + pbase)[col[rr]]);
                    ol1[rr] = lo_decode(o1[rr], (m.ol1 + pbase)[col[rr]]);
                    ol2[rr] = lo_decode(o2[rr], (m.ol2 + pbase)[col[rr]]);
                    // lo parts of the current level, decoded against the
                    // stored hi of plane i
                    cl0[rr] = lo_decode((m.c0 + pbase)[col[rr]], (m.cl0 + pbase)[col[rr]]);
                    cl1[rr] = lo_decode((m.c1 + pbase)[col[rr]], (m.cl1 + pbase)[col[rr]]);
                    cl2[rr] = (i == 0 && k == 0) ? R(0)
                                                 : lo_decode((m.c2 + pbase)[col[rr]], (m.cl2 + pbase)[col[rr]]);
                }
            }
        }

        // normal stresses at plane ip, rows 1..NR+1 (base_solver.py:329-351, 408-418)
        R n1_1[NRR], n2_1[NRR], n3_1[NRR];
        n1_1[0] = n2_1[0] = n3_1[0] = R(0);
        const bool ipn = ip >= 1 && ip <= m.nx - 2;
        const R hsx1 = ipn ? m.sx[ip - 1] : R(1);
#pragma unroll
        for (int rr = 1; rr < NRR; ++rr) {
            R a = 0, b = 0, c = 0;
            const R uz1_km = lane_below(z1[rr]);
            if (ipn && rjn[rr] && kf) {
                const R* cm = smat + MAT_STRIDE * d1[rr];
                const R dux = x1[rr] - x0[rr], duy = y1[rr] - y1[rr - 1];
                if (k >= 1) normal3<EXACT>(cm, dux, duy, z1[rr] - uz1_km, hsx1, hsy[rr], hsz, a, b, c);
                else normal3_surface<EXACT>(cm, dux, duy, z1[rr], m.sx0, m.sy0, m.sz0, a, b, c);
            }
            n1_1[rr] = a;
            n2_1[rr] = b;
            n3_1[rr] = c;
        }
        // shear stresses at plane i (base_solver.py:353-372, 420-433)
        R s4[NRR], s5[NRR], s6[NRR];
        s4[NRR - 1] = s5[NRR - 1] = s6[NRR - 1] = s5[0] = R(0);
        const bool ii = i >= 0 && i <= m.nx - 2;
        const R hfx = ii ? m.fx[i] : R(1);
#pragma unroll
        for (int rr = 0; rr <= NR; ++rr) {
            const R uy0_kp = lane_above(y0[rr]);
            const int d0_jkp = dpp_from_above(d0[rr + 1]);
            R a = 0;
            if (ii && i >= 1 && rjf[rr] && kf) {
                if (k >= 1) a = shear<EXACT>(smat[MAT_STRIDE * d0_jkp + M_C33], uy0_kp - y0[rr], hfz,
                                             z0[rr + 1] - z0[rr], hfy[rr]);
                else a = shear<EXACT>(smat[MAT_STRIDE * d0[rr + 1] + M_C33], uy0_kp - y0[rr], m.fy0,
                                      z0[rr + 1] - z0[rr], m.fz0);
            }
            s4[rr] = a;
            R c = 0;
            if (ii && rjf[rr] && kf) {
                const R cm = smat[MAT_STRIDE * d1[rr + 1] + M_C55];
                if (k >= 1) c = shear<EXACT>(cm, x0[rr + 1] - x0[rr], hfy[rr], y1[rr] - y0[rr], hfx);
                else c = shear<EXACT>(cm, x0[rr + 1] - x0[rr], m.fx0, y1[rr] - y0[rr], m.fz0);
            }
            s6[rr] = c;
        }
#pragma unroll
        for (int rr = 1; rr <= NR; ++rr) {
            const R ux0_kp = lane_above(x0[rr]);
            const int d1_kp = dpp_from_above(d1[rr]);
            R b = 0;
            if (ii && rjn[rr] && kf) {
                if (k >= 1) b = shear<EXACT>(smat[MAT_STRIDE * d1_kp + M_C44], ux0_kp - x0[rr], hfz,
                                             z1[rr] - z0[rr], hfx);
                else b = shear<EXACT>(smat[MAT_STRIDE * d1[rr] + M_C44], ux0_kp - x0[rr], m.fx0,
                                      z1[rr] - z0[rr], m.fz0);
            }
            s5[rr] = b;
        }

        // stage-I update of plane i, output rows 1..NR
        const bool inx = i >= 1 && i <= m.nx - 2;
        const R hsx = inx ? m.sx[i - 1] : R(1);
#pragma unroll
        for (int rr = 1; rr <= NR; ++rr) {
            const int j = j0 - 1 + rr;
            const R n3_kp = lane_above(n3_0[rr]);
            const R s4_km = lane_below(s4[rr]);
            const R s5_km = lane_below(s5[rr]);
            const int d0_kp = dpp_from_above(d0[rr]);
            if (!(outp && rv[rr] && klane)) continue;
            const R s = SPONGE ? damp_at(m, i, j, k) : R(0);
            // ux (base_solver.py:441-447, 495-501)
            if (i <= m.nx - 3 && rjn[rr] && k <= m.nz - 2) {
                R dv;
                if (k >= 1) dv = div3<EXACT>(n1_1[rr] - n1_0[rr], hfx, s6[rr] - s6[rr - 1], hsy[rr], s5[rr] - s5_km, hsz);
                else dv = div3<EXACT>(n1_1[rr] - n1_0[rr], m.fx0, s6[rr] - s6[rr - 1], m.sy0, s5[rr] - R(0), m.sz0);
                const V<R> cur = {x0[rr], cl0[rr]};
                put_with_zface<0, MODE>(m, i, j, k,
                                        leap_from<MODE, SPONGE>(cur, o0[rr], ol0[rr], smat[MAT_STRIDE * d1[rr] + M_COEF], dv, s),
                                        cur, w);
            }
            // uy (base_solver.py:449-455, 503-509)
            if (i <= m.nx - 2 && j >= 1 && j <= m.ny - 3 && k <= m.nz - 2) {
                const V<R> cur = {y0[rr], cl1[rr]};
                V<R> s0 = {R(0), R(0)};
                if (i > 0) {
                    R dv;
                    if (k >= 1) dv = div3<EXACT>(s6[rr] - s6m[rr], hsx, n2_0[rr + 1] - n2_0[rr], hfy[rr], s4[rr] - s4_km, hsz);
                    else dv = div3<EXACT>(s6[rr] - s6m[rr], m.sx0, n2_0[rr + 1] - n2_0[rr], m.fy0, s4[rr] - R(0), m.sz0);
                    s0 = leap_from<MODE, SPONGE>(cur, o1[rr], ol1[rr], smat[MAT_STRIDE * d0[rr + 1] + M_COEF], dv, s);
                }
                put_with_zface<1, MODE>(m, i, j, k, s0, cur, w);
            }
            // uz (base_solver.py:457-463, 511-517)
            if (i <= m.nx - 2 && rjn[rr] && k <= m.nz - 3) {
                const V<R> cur = {z0[rr], cl2[rr]};
                V<R> s0 = {R(0), R(0)};
                if (i > 0) {
                    R dv, coef;
                    if (k >= 1) {
                        dv = div3<EXACT>(s5[rr] - s5m[rr], hsx, s4[rr] - s4[rr - 1], hsy[rr], n3_kp - n3_0[rr], hfz);
                        coef = smat[MAT_STRIDE * d0_kp + M_COEF];
                    } else {
                        dv = m.surf ? div3<EXACT>(s5[rr] - s5m[rr], m.sx0, s4[rr] - s4[rr - 1], m.sy0,
                                                  n3_kp - n3_0[rr], m.fz0)
                                    : div3_uz_surface<EXACT>(s5[rr] - s5m[rr], m.sx0, s4[rr] - s4[rr - 1], m.sy0,
                                                             n3_kp, n3_0[rr], m.fz0);
                        coef = smat[MAT_STRIDE * d0[rr] + M_COEF];
                    }
                    s0 = leap_from<MODE, SPONGE>(cur, o2[rr], ol2[rr], coef, dv, s);
                }
                put_with_zface<2, MODE>(m, i, j, k, s0, cur, w);
            }
        }
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) {
            x0[rr] = x1[rr];
            y0[rr] = y1[rr];
            z0[rr] = z1[rr];
            d0[rr] = d1[rr];
            n1_0[rr] = n1_1[rr];
            n2_0[rr] = n2_1[rr];
            n3_0[rr] = n3_1[rr];
            s5m[rr] = s5[rr];
            s6m[rr] = s6[rr];
        }
    }
}

inline StripMap strip_map(const pf_ctx* c, int nr, int wv, int xchunk) {
    StripMap sm;
    sm.nkt = (c->nz + KOUT - 1) / KOUT;
    const int nstrips = (c->ny + nr - 1) / nr;
    sm.nsg = (nstrips + wv - 1) / wv;
    if (xchunk <= 0) {
        const long long per = (long long)sm.nkt * sm.nsg;
        long long want = (4096 + per - 1) / per;
        if (want < 1) want = 1;
        xchunk = (int)((c->nx + want - 1) / want);
        if (xchunk < 16) xchunk = 16;
    }
    if (xchunk > c->nx) xchunk = c->nx;
    sm.xchunk = xchunk;
    sm.nxc = (c->nx + xchunk - 1) / xchunk;
    return sm;
}

// Rows per strip are the context's tile_j (4 or 8); 4 strips per block.
template <class R, int MODE, bool SPONGE>
int launch_strip_s(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    constexpr int WV = 4;
    const int nr = c->tile_j == 8 ? 8 : 4;
    StripMap sm = strip_map(c, nr, WV, c->tile_x);
    const long long nb = (long long)sm.nkt * sm.nsg * sm.nxc;
    if (nb > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many tiles");
    const dim3 grd((unsigned)nb), blk(64 * WV);
    if (nr == 8) hipLaunchKernelGGL((k_sweep_strip<R, MODE, SPONGE, 8, WV>), grd, blk, 0, s, m, sm);
    else hipLaunchKernelGGL((k_sweep_strip<R, MODE, SPONGE, 4, WV>), grd, blk, 0, s, m, sm);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

template <class R, int MODE>
int launch_strip(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    return c->sponge ? launch_strip_s<R, MODE, true>(c, m, s) : launch_strip_s<R, MODE, false>(c, m, s);
}

}  // namespace pf
