// Material map from an inclusion list, on the device: the reference's
// Grid.inclusionIndices + Material.setConstants (grid.py:158-175,
// material.py:55-63), which loop over the inclusions in Python with one
// (ny, nx) meshgrid per inclusion and a fancy-index write of the whole
// (nx, ny, nz) map.
//
// Per inclusion m (a vertical cylinder: centre tx, ty, radius r, as the
// reference's float32 target record holds them) a node column (i, j) is
// inside when sqrt((x[i]-tx)^2 + (y[j]-ty)^2) < r, evaluated in FP64 in the
// reference's order (X = x - tx, Y = y - ty, then sqrt(X*X + Y*Y); the
// library is built with -ffp-contract=off, so no product is fused); the
// inclusion covers the column's first K_m nodes along z (the host derives
// K_m from the z test, including the reference's reuse of the previous
// target's index array).  A node is secondary (index 1) when any inclusion
// covers it, else primary (0).
//
// k_raster_cols: one thread per column, the largest K_m of the inclusions
// that contain it.  k_raster_fill: one thread per 4 nodes along z, the map
// bytes in the field layout.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pf {

struct RasterTarget {
    double x, y, r, k;  // centre, radius, covered z nodes
};

__global__ __launch_bounds__(256) void k_raster_cols(const double* __restrict__ x, const double* __restrict__ y, int nx,
                                                     int ny, const RasterTarget* __restrict__ tg, int nt,
                                                     int* __restrict__ kmax) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nx * ny) return;
    const int i = (int)(t / ny), j = (int)(t % ny);
    const double xi = x[i], yj = y[j];
    int km = 0;
    for (int m = 0; m < nt; m++) {
        const RasterTarget g = tg[m];
        const double X = xi - g.x;
        const double Y = yj - g.y;
        const double R = __builtin_sqrt(X * X + Y * Y);
        if (R < g.r && (int)g.k > km) km = (int)g.k;
    }
    kmax[t] = km;
}

__global__ __launch_bounds__(256) void k_raster_fill(const int* __restrict__ kmax, long long ncols, int nz,
                                                     uint8_t* __restrict__ ids) {
    // 4 nodes per thread along z; a column's bytes are contiguous
    const int per = (nz + 3) / 4;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncols * per) return;
    const long long c = t / per;
    const int k0 = (int)(t % per) * 4;
    const int km = kmax[c];
    uint8_t* out = ids + c * nz;
    for (int k = k0; k < k0 + 4 && k < nz; k++) out[k] = k < km ? 1 : 0;
}

}  // namespace pf
