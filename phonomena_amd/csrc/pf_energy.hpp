// Energy audit kernels: E = dV/2 [ sum rho_q ((u_q - u_q_old)/dt)^2
//                                 + sum over stress sites of T(u) : strain(u_old) ],
// the quantity the leapfrog scheme conserves.
// Evaluated in FP64 on an FP64 view of the two newest levels; the stresses
// and strain brackets are exactly those of the exact stencil formulas.
// Per-thread sums, a wavefront/LDS tree per block, then one block adds the
// block partials in a fixed order, so the result is deterministic.
#pragma once
#include "pf_kernels.hpp"

namespace pf {

struct EnergyView {
    Mesh<double> m;            // c* = newest level, n* = previous level (read only)
    const double* rho;         // density per material
    double dt;
    int subst_prev;            // previous level's uz wall line holds the source
    double w_prev;             // value of the step that produced the newest level
};

PF_DEV double warp_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Leapfrog-invariant potential term at one node site: the stress of the
// newest level (c* arrays) times the strain bracket of the previous level
// (n* arrays), over the reference's index ranges.  As in the reference's
// arrays after time_step, the previous level carries the source value on
// its uz wall line (u_old was copied after the source write) and the newest
// level carries what the step stored there.
PF_DEV double site_energy(const EnergyView& ev, int i, int j, int k) {
    const Mesh<double>& m = ev.m;
    double acc = 0.0;
    auto C0 = [&](int a, int b, int c) { return m.c0[at(m, a, b, c)]; };
    auto C1 = [&](int a, int b, int c) { return m.c1[at(m, a, b, c)]; };
    auto C2 = [&](int a, int b, int c) { return m.c2[at(m, a, b, c)]; };
    auto N0 = [&](int a, int b, int c) { return (double)m.n0[at(m, a, b, c)]; };
    auto N1 = [&](int a, int b, int c) { return (double)m.n1[at(m, a, b, c)]; };
    auto N2 = [&](int a, int b, int c) {
        return (ev.subst_prev && a == 0 && c == 0) ? ev.w_prev : (double)m.n2[at(m, a, b, c)];
    };
    // normal sites
    if (i >= 1 && i <= m.nx - 2 && j >= 1 && j <= m.ny - 2 && k <= m.nz - 2) {
        const double* c = matp(m, i, j, k);
        double dux = C0(i, j, k) - C0(i - 1, j, k), oux = N0(i, j, k) - N0(i - 1, j, k);
        double duy = C1(i, j, k) - C1(i, j - 1, k), ouy = N1(i, j, k) - N1(i, j - 1, k);
        double t1, t2, t3, e1, e2, e3;
        if (k >= 1) {
            double duz = C2(i, j, k) - C2(i, j, k - 1), ouz = N2(i, j, k) - N2(i, j, k - 1);
            normal3<true>(c, dux, duy, duz, m.sx[i - 1], m.sy[j - 1], m.sz[k - 1], t1, t2, t3);
            e1 = oux / m.sx[i - 1];
            e2 = ouy / m.sy[j - 1];
            e3 = ouz / m.sz[k - 1];
        } else {
            normal3_surface<true>(c, dux, duy, C2(i, j, 0), m.sx0, m.sy0, m.sz0, t1, t2, t3);
            e1 = oux / m.sx0;
            e2 = ouy / m.sy0;
            e3 = (N2(i, j, 0) - 0.0) / m.sz0;
        }
        acc += t1 * e1 + t2 * e2 + t3 * e3;
    }
    // T4 sites
    if (i >= 1 && i <= m.nx - 2 && j <= m.ny - 2 && k <= m.nz - 2) {
        double g, go, c;
        if (k >= 1) {
            c = matp(m, i, j + 1, k + 1)[M_C33];
            g = (C1(i, j, k + 1) - C1(i, j, k)) / m.fz[k] + (C2(i, j + 1, k) - C2(i, j, k)) / m.fy[j];
            go = (N1(i, j, k + 1) - N1(i, j, k)) / m.fz[k] + (N2(i, j + 1, k) - N2(i, j, k)) / m.fy[j];
        } else {
            c = matp(m, i, j + 1, 0)[M_C33];
            g = (C1(i, j, 1) - C1(i, j, 0)) / m.fy0 + (C2(i, j + 1, 0) - C2(i, j, 0)) / m.fz0;
            go = (N1(i, j, 1) - N1(i, j, 0)) / m.fy0 + (N2(i, j + 1, 0) - N2(i, j, 0)) / m.fz0;
        }
        acc += (c * g) * go;
    }
    // T5 sites
    if (i <= m.nx - 2 && j >= 1 && j <= m.ny - 2 && k <= m.nz - 2) {
        double g, go, c;
        if (k >= 1) {
            c = matp(m, i + 1, j, k + 1)[M_C44];
            g = (C0(i, j, k + 1) - C0(i, j, k)) / m.fz[k] + (C2(i + 1, j, k) - C2(i, j, k)) / m.fx[i];
            go = (N0(i, j, k + 1) - N0(i, j, k)) / m.fz[k] + (N2(i + 1, j, k) - N2(i, j, k)) / m.fx[i];
        } else {
            c = matp(m, i + 1, j, 0)[M_C44];
            g = (C0(i, j, 1) - C0(i, j, 0)) / m.fx0 + (C2(i + 1, j, 0) - C2(i, j, 0)) / m.fz0;
            go = (N0(i, j, 1) - N0(i, j, 0)) / m.fx0 + (N2(i + 1, j, 0) - N2(i, j, 0)) / m.fz0;
        }
        acc += (c * g) * go;
    }
    // T6 sites
    if (i <= m.nx - 2 && j <= m.ny - 2 && k <= m.nz - 2) {
        double g, go, c;
        if (k >= 1) {
            c = matp(m, i + 1, j + 1, k)[M_C55];
            g = (C0(i, j + 1, k) - C0(i, j, k)) / m.fy[j] + (C1(i + 1, j, k) - C1(i, j, k)) / m.fx[i];
            go = (N0(i, j + 1, k) - N0(i, j, k)) / m.fy[j] + (N1(i + 1, j, k) - N1(i, j, k)) / m.fx[i];
        } else {
            c = matp(m, i + 1, j + 1, 0)[M_C55];
            g = (C0(i, j + 1, 0) - C0(i, j, 0)) / m.fx0 + (C1(i + 1, j, 0) - C1(i, j, 0)) / m.fz0;
            go = (N0(i, j + 1, 0) - N0(i, j, 0)) / m.fx0 + (N1(i + 1, j, 0) - N1(i, j, 0)) / m.fz0;
        }
        acc += (c * g) * go;
    }
    return acc;
}

PF_DEV double node_kinetic(const EnergyView& ev, int i, int j, int k) {
    const Mesh<double>& m = ev.m;
    double acc = 0.0;
    long long id = at(m, i, j, k);
    if (i <= m.nx - 2) {
        double v = (m.c0[id] - m.n0[id]) / ev.dt;
        acc += ev.rho[m.ids[at(m, i + 1, j, k)]] * (v * v);
    }
    if (i >= 1 && j <= m.ny - 2) {  // x = 0 wall of uy is clamped: excluded
        double v = (m.c1[id] - m.n1[id]) / ev.dt;
        acc += ev.rho[m.ids[at(m, i, j + 1, k)]] * (v * v);
    }
    if (i >= 1 && k <= m.nz - 2) {  // x = 0 wall of uz is driven: excluded
        double v = (m.c2[id] - m.n2[id]) / ev.dt;
        acc += ev.rho[m.ids[at(m, i, j, k + 1)]] * (v * v);
    }
    return acc;
}

static __global__ __launch_bounds__(256) void k_energy_partial(EnergyView ev, double* __restrict__ partials) {
    const Mesh<double>& m = ev.m;
    __shared__ double red[2][4];
    double ke = 0.0, se = 0.0;
    const long long n = (long long)m.nx * m.ny * m.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int k = (int)(t % m.nz);
        long long r = t / m.nz;
        int j = (int)(r % m.ny);
        int i = (int)(r / m.ny);
        ke += node_kinetic(ev, i, j, k);
        se += site_energy(ev, i, j, k);
    }
    ke = warp_sum(ke);
    se = warp_sum(se);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[0][wv] = ke; red[1][wv] = se; }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        partials[2 * blockIdx.x + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

static __global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ partials, int nblk,
                                                      double* __restrict__ out) {
    __shared__ double red[2][256];
    double ke = 0.0, se = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        ke += partials[2 * b];
        se += partials[2 * b + 1];
    }
    red[0][threadIdx.x] = ke;
    red[1][threadIdx.x] = se;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = red[0][0]; out[1] = red[1][0]; }
}

}  // namespace pf
