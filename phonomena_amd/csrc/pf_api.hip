// C ABI of libphonofdtd (declared in include/phonofdtd.h): context life
// cycle, stepping, field transfer, probes, energy audit, profiling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/phonofdtd.h"
#include "pf_energy.hpp"
#include "pf_kernels.hpp"
#include "pf_runtime.hpp"
#include "pf_strip.hpp"
#include "pf_fused.hpp"
#include "pf_raster.hpp"
#include "pf_tiled.hpp"

using namespace pf;

// ------------------------------------------------------------ errors ----
static thread_local char g_err[512] = "";

int pf::fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* pf_last_error(void) { return g_err; }
extern "C" int pf_abi_version(void) { return PF_ABI_VERSION; }

extern "C" int pf_device_info(char* name, int32_t cap, int32_t* cu_count, double* hbm_gib) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    if (name && cap > 0) {
        snprintf(name, (size_t)cap, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_gib) *hbm_gib = (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    return PF_OK;
}

// ----------------------------------------------------------- helpers ----
namespace {

template <class R>
int upload_table(void** dst, const std::vector<R>& v) {
    if (v.empty()) { *dst = nullptr; return PF_OK; }
    HIPCHK(hipMalloc(dst, v.size() * sizeof(R)));
    HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(R), hipMemcpyHostToDevice));
    return PF_OK;
}

template <class R>
std::vector<R> spacing_table(const double* src, int n, bool invert) {
    std::vector<R> v((size_t)n);
    for (int i = 0; i < n; i++) v[(size_t)i] = invert ? (R)(1.0 / src[i]) : (R)src[i];
    return v;
}

// Material rows of 16: C00..C22, C33, C44, C55, dt2/rho (as the host folds
// it, base_solver.py:443 `self.m.dt**2/self.m.P[...]`), padding.
template <class R>
std::vector<R> material_rows(const double* m13, int n, double dt2) {
    std::vector<R> v((size_t)n * MAT_STRIDE, (R)0);
    for (int r = 0; r < n; r++) {
        for (int e = 0; e < 12; e++) v[(size_t)r * MAT_STRIDE + e] = (R)m13[r * 13 + e];
        v[(size_t)r * MAT_STRIDE + M_COEF] = (R)(dt2 / m13[r * 13 + 12]);
    }
    return v;
}

int check_desc(const pf_desc* d) {
    if (!d) return fail(PF_ERR_ARG, "desc is NULL");
    if (d->nx < 4 || d->ny < 4 || d->nz < 4)
        return fail(PF_ERR_ARG, "node counts must be >= 4 per axis (got %d x %d x %d)", d->nx, d->ny, d->nz);
    if ((long long)d->nx * d->ny * d->nz > (1LL << 40)) return fail(PF_ERR_ARG, "mesh too large");
    if (d->precision < PF_FP64 || d->precision > PF_FP32C) return fail(PF_ERR_ARG, "bad precision %d", d->precision);
    if (d->kernel < PF_KERNEL_AUTO || d->kernel > PF_KERNEL_TILED2) return fail(PF_ERR_ARG, "bad kernel %d", d->kernel);
    if (d->n_materials < 1 || d->n_materials > 256) return fail(PF_ERR_ARG, "n_materials must be 1..256");
    if (!d->fdx || !d->fdy || !d->fdz || !d->sdx || !d->sdy || !d->sdz) return fail(PF_ERR_ARG, "spacing table missing");
    if (!d->materials) return fail(PF_ERR_ARG, "material table missing");
    if (!d->ids) {
        if (!d->node_x || !d->node_y || d->n_targets < 0 || (d->n_targets > 0 && !d->targets))
            return fail(PF_ERR_ARG, "ids missing and no inclusion list to rasterize");
        if (d->n_materials < 2 && d->n_targets > 0)
            return fail(PF_ERR_ARG, "inclusions need a secondary material (n_materials >= 2)");
        for (int m = 0; m < d->n_targets; m++) {
            const double k = d->targets[4 * m + 3];
            if (!(k >= 0 && k <= d->nz && k == (double)(int)k))
                return fail(PF_ERR_ARG, "target %d covers %g z nodes (must be an integer in 0..nz)", m, k);
        }
    }
    if (!d->mur) return fail(PF_ERR_ARG, "Mur coefficients missing");
    if (d->surface != 0 && d->surface != 1) return fail(PF_ERR_ARG, "surface must be 0 or 1 (got %d)", d->surface);
    if (!d->wave || d->wave_len < 0) return fail(PF_ERR_ARG, "source table missing");
    if (!(d->dt > 0) || !(d->dt2 > 0)) return fail(PF_ERR_ARG, "dt must be > 0");
    const bool any = d->sponge_x || d->sponge_y || d->sponge_z;
    const bool all = d->sponge_x && d->sponge_y && d->sponge_z;
    if (any && !all) return fail(PF_ERR_ARG, "sponge needs all three axis tables or none");
    for (int i = 0; i < d->nx - 1; i++) if (!(d->fdx[i] > 0)) return fail(PF_ERR_ARG, "fdx[%d] <= 0", i);
    for (int i = 0; i < d->ny - 1; i++) if (!(d->fdy[i] > 0)) return fail(PF_ERR_ARG, "fdy[%d] <= 0", i);
    for (int i = 0; i < d->nz - 1; i++) if (!(d->fdz[i] > 0)) return fail(PF_ERR_ARG, "fdz[%d] <= 0", i);
    for (int r = 0; r < d->n_materials; r++)
        if (!(d->materials[r * 13 + 12] > 0)) return fail(PF_ERR_ARG, "material %d has rho <= 0", r);
    const long long nn = (long long)d->nx * d->ny * d->nz;
    if (d->ids)
        for (long long t = 0; t < nn; t++)
            if (d->ids[t] >= d->n_materials)
                return fail(PF_ERR_ARG, "ids[%lld] = %d out of range", t, (int)d->ids[t]);
    if (d->probe[2] > 0) {
        const int* p = d->probe;
        if (p[1] < 1 || p[4] < 1 || p[5] < 1 || p[0] < 0 || p[3] < 0)
            return fail(PF_ERR_ARG, "bad probe grid");
        if (p[0] + p[1] * (p[2] - 1) > d->nx - 1 || p[3] + p[4] * (p[5] - 1) > d->ny - 1)
            return fail(PF_ERR_ARG, "probe grid outside the u_z surface");
        if (d->record_every < 1 || d->probe_capacity < 1) return fail(PF_ERR_ARG, "record_every and probe_capacity must be >= 1");
    }
    return PF_OK;
}

}  // namespace

// --------------------------------------------------------- life cycle ----
// Whether every spacing is a power of two whose reciprocal is a normal
// double: then 1/h is exact and a product with it is bit-identical to the
// division by h.
static bool pow2_spacings(const pf_desc* d) {
    const double* t[6] = {d->fdx, d->fdy, d->fdz, d->sdx, d->sdy, d->sdz};
    const int n[6] = {d->nx - 1, d->ny - 1, d->nz - 1, d->nx - 2, d->ny - 2, d->nz - 2};
    for (int a = 0; a < 6; a++)
        for (int i = 0; i < n[a]; i++) {
            int e = 0;
            const double f = std::frexp(t[a][i], &e);
            if (!std::isfinite(t[a][i]) || f != 0.5 || e < -1000 || e > 1000) return false;
        }
    return true;
}

template <class R>
static int create_tables(pf_ctx* c, const pf_desc* d) {
    const bool exact = d->precision == PF_FP64;
    int rc;
    // the FP64 one-step tiled sweep on power-of-two spacings multiplies by
    // exact reciprocals (MODE_FP64P)
    c->recip = exact && c->kernel == PF_KERNEL_TILED && pow2_spacings(d);
    if (c->recip) {
        if ((rc = upload_table(&c->rfx, spacing_table<double>(d->fdx, d->nx - 1, true)))) return rc;
        if ((rc = upload_table(&c->rfy, spacing_table<double>(d->fdy, d->ny - 1, true)))) return rc;
        if ((rc = upload_table(&c->rfz, spacing_table<double>(d->fdz, d->nz - 1, true)))) return rc;
        if ((rc = upload_table(&c->rsx, spacing_table<double>(d->sdx, d->nx - 2, true)))) return rc;
        if ((rc = upload_table(&c->rsy, spacing_table<double>(d->sdy, d->ny - 2, true)))) return rc;
        if ((rc = upload_table(&c->rsz, spacing_table<double>(d->sdz, d->nz - 2, true)))) return rc;
    }
    if ((rc = upload_table(&c->fx, spacing_table<R>(d->fdx, d->nx - 1, !exact)))) return rc;
    if ((rc = upload_table(&c->fy, spacing_table<R>(d->fdy, d->ny - 1, !exact)))) return rc;
    if ((rc = upload_table(&c->fz, spacing_table<R>(d->fdz, d->nz - 1, !exact)))) return rc;
    if ((rc = upload_table(&c->sx, spacing_table<R>(d->sdx, d->nx - 2, !exact)))) return rc;
    if ((rc = upload_table(&c->sy, spacing_table<R>(d->sdy, d->ny - 2, !exact)))) return rc;
    if ((rc = upload_table(&c->sz, spacing_table<R>(d->sdz, d->nz - 2, !exact)))) return rc;
    if ((rc = upload_table(&c->mat, material_rows<R>(d->materials, d->n_materials, d->dt2)))) return rc;
    if (c->sponge) {
        if ((rc = upload_table(&c->dpx, spacing_table<R>(d->sponge_x, d->nx, false)))) return rc;
        if ((rc = upload_table(&c->dpy, spacing_table<R>(d->sponge_y, d->ny, false)))) return rc;
        if ((rc = upload_table(&c->dpz, spacing_table<R>(d->sponge_z, d->nz, false)))) return rc;
    }
    return PF_OK;
}

static int tune_shape(pf_ctx* c);

extern "C" int pf_create(const pf_desc* d, pf_ctx** out) {
    if (!out) return fail(PF_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_desc(d);
    if (rc) return rc;
    pf_ctx* c = new (std::nothrow) pf_ctx();
    if (!c) return fail(PF_ERR_NOMEM, "host allocation failed");
    rc = ctx_init(c, d);
    if (!rc && c->shape_open) rc = tune_shape(c);
    if (rc) {
        ctx_free(c);
        delete c;
        return rc;
    }
    *out = c;
    return PF_OK;
}

// Material map of the context from the descriptor's inclusion list.
static int rasterize(pf_ctx* c, const pf_desc* d) {
    double *x = nullptr, *y = nullptr;
    RasterTarget* tg = nullptr;
    int* kmax = nullptr;
    const long long ncols = (long long)d->nx * d->ny;
    const int nt = d->n_targets;
    auto cleanup = [&]() {
        if (x) (void)hipFree(x);
        if (y) (void)hipFree(y);
        if (tg) (void)hipFree(tg);
        if (kmax) (void)hipFree(kmax);
    };
    int rc = PF_OK;
    do {
        std::vector<RasterTarget> h((size_t)(nt > 0 ? nt : 1));
        for (int m = 0; m < nt; m++)
            h[(size_t)m] = {d->targets[4 * m], d->targets[4 * m + 1], d->targets[4 * m + 2], d->targets[4 * m + 3]};
        hipError_t e;
        if ((e = hipMalloc((void**)&x, (size_t)d->nx * sizeof(double))) != hipSuccess ||
            (e = hipMalloc((void**)&y, (size_t)d->ny * sizeof(double))) != hipSuccess ||
            (e = hipMalloc((void**)&tg, h.size() * sizeof(RasterTarget))) != hipSuccess ||
            (e = hipMalloc((void**)&kmax, (size_t)ncols * sizeof(int))) != hipSuccess ||
            (e = hipMemcpy(x, d->node_x, (size_t)d->nx * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(y, d->node_y, (size_t)d->ny * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(tg, h.data(), h.size() * sizeof(RasterTarget), hipMemcpyHostToDevice)) != hipSuccess) {
            rc = fail(PF_ERR_HIP, "rasterize: %s", hipGetErrorString(e));
            break;
        }
        hipLaunchKernelGGL(k_raster_cols, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, c->stream, x, y, d->nx,
                           d->ny, tg, nt, kmax);
        const long long n4 = ncols * ((d->nz + 3) / 4);
        hipLaunchKernelGGL(k_raster_fill, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, c->stream, kmax, ncols,
                           d->nz, c->ids);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(c->stream)) != hipSuccess)
            rc = fail(PF_ERR_HIP, "rasterize: %s", hipGetErrorString(e));
    } while (0);
    cleanup();
    return rc;
}

int pf::ctx_init(pf_ctx* c, const pf_desc* d) {
    c->nx = d->nx;
    c->ny = d->ny;
    c->nz = d->nz;
    // rows are not padded: with z contiguous the sweep streams whole planes,
    // and padding rows to a cache-line multiple would only add traffic
    c->pz = d->nz;
    c->plane = (long long)d->ny * c->pz;
    c->cells = (long long)d->nx * c->plane;
    c->precision = d->precision;
    const bool few_mats = d->n_materials <= PF_TILED_MAX_MATERIALS;
    // the tiled sweep addresses a whole level with 32-bit offsets
    const bool small = level_bytes(c->cells, d->precision) < (1ULL << 32);
    // two-step sweep: measured faster than the one-step sweep in FP32 from
    // about 2^24 nodes on (fewer passes over memory outweigh its band and
    // second face launch); FP32C runs it out of registers
    const bool pair_auto = d->precision == PF_FP32 && d->nx >= 8 && d->ny >= 8 && d->nz >= 8 &&
                           (long long)d->nx * d->ny * d->nz >= PF_TILED2_AUTO_NODES;
    c->kernel = d->kernel != PF_KERNEL_AUTO ? d->kernel
                : !(few_mats && small)      ? PF_KERNEL_NAIVE
                : pair_auto                 ? PF_KERNEL_TILED2
                                            : PF_KERNEL_TILED;
    if (c->kernel != PF_KERNEL_NAIVE && !few_mats)
        return fail(PF_ERR_ARG, "the tiled and strip sweeps hold at most %d materials (got %d)",
                    PF_TILED_MAX_MATERIALS, d->n_materials);
    c->tile_x = d->tile_x > 0 ? d->tile_x : 0;
    if (c->kernel == PF_KERNEL_TILED && !small)
        return fail(PF_ERR_ARG, "the tiled sweep needs a time level < 4 GiB (use the naive kernel)");
    const bool pair = c->kernel == PF_KERNEL_TILED2;
    if (pair) {
        if (d->precision == PF_FP64)
            return fail(PF_ERR_ARG, "the two-step sweep is built for PF_FP32 and PF_FP32C");
        if (d->nx < 8 || d->ny < 8 || d->nz < 8)
            return fail(PF_ERR_ARG, "the two-step sweep needs >= 8 nodes per axis (got %d x %d x %d)", d->nx, d->ny,
                        d->nz);
        if (!small) return fail(PF_ERR_ARG, "the two-step sweep needs a time level < 4 GiB");
        if (c->tile_x > XT2 - 6) return fail(PF_ERR_ARG, "tile_x must be <= %d (got %d)", XT2 - 6, c->tile_x);
        // 16 rows of 64 lanes or 32 rows of 32 lanes (1024 threads); by
        // default the one with fewer idle lane-slots for this mesh
        int kz = d->tile_z, tj = d->tile_j;
        if (kz == 0 && tj != 0) kz = tj == 32 ? 32 : 64;
        if (kz == 0) {
            const TileMap a = fused_map(c, 16, 64, c->tile_x), b = fused_map(c, 32, 32, c->tile_x);
            const double sa = (double)a.nkt * 64 * a.njt * 16 * a.nxc * (a.xchunk + 3);
            const double sb = (double)b.nkt * 32 * b.njt * 32 * b.nxc * (b.xchunk + 3);
            kz = sb <= sa ? 32 : 64;
        }
        if (tj == 0) tj = kz == 64 ? 16 : 32;
        if (!((kz == 64 && tj == 16) || (kz == 32 && tj == 32)))
            return fail(PF_ERR_ARG, "two-step tiles are 16 x 64 or 32 x 32 (got tile_j %d, tile_z %d)", tj, kz);
        c->pair_j = tj;
        c->pair_z = kz;
    }
    if (c->kernel == PF_KERNEL_STRIP) {
        c->tile_j = d->tile_j > 0 ? d->tile_j : 4;
        if (c->tile_j != 4 && c->tile_j != 8)
            return fail(PF_ERR_ARG, "strip rows must be 4 or 8 (got %d)", c->tile_j);
    } else {
        // Default shapes, from measurement on the 513x513x257 crystal: FP32
        // runs best with 64-lane rows and 16 rows (1024 threads), FP64 (whose
        // registers allow 512 threads) with 32-lane rows and 8 rows.  The
        // other layout is taken when it leaves 15% fewer lanes idle.
        const bool f64 = d->precision == PF_FP64;
        const int tj64 = f64 ? 8 : 16, tj32 = f64 ? 8 : 32;
        // the one-step sweep of a two-step context (odd steps) takes its
        // default shape
        int kz = pair ? 0 : d->tile_z;
        const int req_j = pair ? 0 : d->tile_j;
        if (kz != 0 && kz != 32 && kz != 64) return fail(PF_ERR_ARG, "tile_z must be 0, 32 or 64 (got %d)", kz);
        if (kz == 0) {
            if (req_j > 0) {
                kz = req_j == 32 ? 32 : 64;
            } else {
                const double s32 = pf::tiled_slots(c, tj32, 32), s64 = pf::tiled_slots(c, tj64, 64);
                kz = f64 ? (s64 < 0.85 * s32 ? 64 : 32) : (s32 < 0.85 * s64 ? 32 : 64);
            }
        }
        c->tile_z = kz;
        c->tile_j = req_j > 0 ? req_j : (kz == 64 ? tj64 : tj32);
        // measured at creation when the caller leaves the tile fields at 0
        // (for a two-step context: the one-step shape of its odd and single
        // steps)
        c->shape_open = (c->kernel == PF_KERNEL_TILED || pair) && d->tile_j == 0 && d->tile_z == 0;
        const bool ok = kz == 64 ? (c->tile_j == 4 || c->tile_j == 8 || c->tile_j == 16)
                                 : (c->tile_j == 8 || c->tile_j == 16 || c->tile_j == 32);
        if (!ok) return fail(PF_ERR_ARG, "tile_j must be 4, 8 or 16 with tile_z 64, 8, 16 or 32 with tile_z 32 (got %d, %d)",
                             c->tile_j, kz);
    }
    c->esz = d->precision == PF_FP64 ? 8 : 4;
    c->sponge = d->sponge_x != nullptr;
    c->surface = d->surface;
    c->dt = d->dt;
    c->nmat = d->n_materials;
    HIPCHK(hipGetDevice(&c->dev));
    {
        int rc = noalloc_scratch_init(c->dev);
        if (rc) return rc;
    }
    // sizes the x chunks of the tiled sweeps (0 keeps the fixed rule)
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->dev) != hipSuccess) {
        (void)hipGetLastError();
        c->cus = 0;
    }
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    // fields: two levels (four for the two-step sweep) x three components
    // (+ 16-bit lo parts in compensated mode), each level one allocation.
    // Level l starts l * PF_LEVEL_STAGGER bytes into its allocation.  The
    // allocations' bases are aligned alike, so the same node of every level
    // would sit at the same low address bits, and a sweep reads and writes
    // the levels at neighbouring nodes together: their lines compete for the
    // same cache sets.  An odd multiple of 256 B spreads them (measured on
    // the 513x513x257-node crystal: the two-step pass, which streams four
    // levels, moved 7.6 % fewer bytes and ran 1.3-1.5 % faster with 256,
    // 2304, 4352 or 8448 B, not with 1024, 4096 or 4608; the one-step
    // sweeps, two levels, changed by 0.2-0.4 %).
    const size_t bytes = (size_t)c->cells * c->esz;
    const size_t lbytes = (size_t)c->cells * sizeof(lo_t);
    const size_t lvl_bytes = level_bytes(c->cells, c->precision);
    c->nlevels = pair ? 4 : 2;
    for (int l = 0; l < c->nlevels; l++) {
        HIPCHK(hipMalloc(&c->lvlbase[l], lvl_bytes + (size_t)l * PF_LEVEL_STAGGER));
        char* b = (char*)c->lvlbase[l] + (size_t)l * PF_LEVEL_STAGGER;
        HIPCHK(hipMemsetAsync(b, 0, lvl_bytes, c->stream));
        for (int q = 0; q < 3; q++) {
            c->lvl[l][q] = b + q * bytes;
            if (c->precision == PF_FP32C) c->lvlo[l][q] = b + 3 * bytes + q * lbytes;
        }
    }
    c->cur = 0;
    c->old = 1;
    // material index map, in the field layout (rows unpadded: one copy; a
    // row-by-row 2-D copy took ~0.3 s per million nodes)
    HIPCHK(hipMalloc((void**)&c->ids, (size_t)c->cells));
    if (!d->ids) {
        int rc = rasterize(c, d);
        if (rc) return rc;
    } else if (c->pz == d->nz) {
        HIPCHK(hipMemcpy(c->ids, d->ids, (size_t)c->cells, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemsetAsync(c->ids, 0, (size_t)c->cells, c->stream));
        HIPCHK(hipMemcpy2DAsync(c->ids, (size_t)c->pz, d->ids, (size_t)d->nz, (size_t)d->nz,
                                (size_t)d->nx * d->ny, hipMemcpyHostToDevice, c->stream));
    }

    int rc = c->esz == 8 ? create_tables<double>(c, d) : create_tables<float>(c, d);
    if (rc) return rc;

    // FP64 copies for the energy audit
    {
        std::vector<double> m64 = material_rows<double>(d->materials, d->n_materials, d->dt2);
        std::vector<double> rho((size_t)d->n_materials);
        for (int r = 0; r < d->n_materials; r++) rho[(size_t)r] = d->materials[r * 13 + 12];
        if ((rc = upload_table(&c->mat64, m64))) return rc;
        if ((rc = upload_table(&c->rho64, rho))) return rc;
        if ((rc = upload_table(&c->fx64, spacing_table<double>(d->fdx, d->nx - 1, false)))) return rc;
        if ((rc = upload_table(&c->fy64, spacing_table<double>(d->fdy, d->ny - 1, false)))) return rc;
        if ((rc = upload_table(&c->fz64, spacing_table<double>(d->fdz, d->nz - 1, false)))) return rc;
        if ((rc = upload_table(&c->sx64, spacing_table<double>(d->sdx, d->nx - 2, false)))) return rc;
        if ((rc = upload_table(&c->sy64, spacing_table<double>(d->sdy, d->ny - 2, false)))) return rc;
        if ((rc = upload_table(&c->sz64, spacing_table<double>(d->sdz, d->nz - 2, false)))) return rc;
    }
    for (int a = 0; a < 3; a++) {
        const double* f = a == 0 ? d->fdx : (a == 1 ? d->fdy : d->fdz);
        const double* s = a == 0 ? d->sdx : (a == 1 ? d->sdy : d->sdz);
        c->f0[a] = f[0];
        c->s0[a] = s[0];
    }
    double mx = 0, my = 0, mz = 0;
    for (int i = 0; i < d->nx - 1; i++) mx += d->fdx[i];
    for (int i = 0; i < d->ny - 1; i++) my += d->fdy[i];
    for (int i = 0; i < d->nz - 1; i++) mz += d->fdz[i];
    c->dV = (mx / (d->nx - 1)) * (my / (d->ny - 1)) * (mz / (d->nz - 1));
    for (int e = 0; e < 12; e++) c->mur[e] = d->mur[e];

    // source table (FP64 on the device) and step counter
    c->wave_len = d->wave_len;
    if (d->wave_len > 0) {
        HIPCHK(hipMalloc((void**)&c->wave, (size_t)d->wave_len * sizeof(double)));
        HIPCHK(hipMemcpy(c->wave, d->wave, (size_t)d->wave_len * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc((void**)&c->d_step, sizeof(long long)));
    HIPCHK(hipMemsetAsync(c->d_step, 0, sizeof(long long), c->stream));

    // probes
    memcpy(c->probe, d->probe, sizeof c->probe);
    if (d->probe[2] > 0) {
        c->every = d->record_every;
        c->cap = d->probe_capacity;
        const size_t pb = (size_t)c->cap * d->probe[2] * d->probe[5] * sizeof(double);
        HIPCHK(hipMalloc((void**)&c->ring, pb));
        HIPCHK(hipMemsetAsync(c->ring, 0, pb, c->stream));
        // frame count, then the probe launch's done count (an int)
        HIPCHK(hipMalloc((void**)&c->d_frame, 2 * sizeof(long long)));
        HIPCHK(hipMemsetAsync(c->d_frame, 0, 2 * sizeof(long long), c->stream));
    }
    HIPCHK(hipMalloc((void**)&c->e_out, 2 * sizeof(double)));
    c->e_blocks = 1024;
    HIPCHK(hipMalloc((void**)&c->e_partials, 2 * sizeof(double) * (size_t)c->e_blocks));
    // the clears above are ordered on the context's own stream; waiting for
    // it (not the whole device) lets a context be created while other
    // contexts' or a batch's work runs
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

void pf::ctx_free(pf_ctx* c) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    for (int l = 0; l < 4; l++)
        F(c->lvlbase[l]);
    F(c->ids);
    F(c->mat); F(c->fx); F(c->fy); F(c->fz); F(c->sx); F(c->sy); F(c->sz);
    F(c->rfx); F(c->rfy); F(c->rfz); F(c->rsx); F(c->rsy); F(c->rsz);
    F(c->dpx); F(c->dpy); F(c->dpz);
    F(c->mat64); F(c->rho64); F(c->fx64); F(c->fy64); F(c->fz64); F(c->sx64); F(c->sy64); F(c->sz64);
    F(c->wave); F(c->d_step); F(c->ring); F(c->d_frame); F(c->e_out); F(c->e_partials);
    F(c->scratch);
    if (c->snap_stream) (void)hipStreamSynchronize(c->snap_stream);  // no copy in flight
    F(c->snap_dev);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    for (int i = 0; i < c->snap_slots; i++) {
        if (c->snap_packed[i]) (void)hipEventDestroy(c->snap_packed[i]);
        if (c->snap_copied[i]) (void)hipEventDestroy(c->snap_copied[i]);
    }
    if (c->snap_stream) (void)hipStreamDestroy(c->snap_stream);
    for (auto& e : c->prof_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

extern "C" int pf_destroy(pf_ctx* c) {
    if (!c) return PF_OK;
    if (c->batch) return fail(PF_ERR_STATE, "context is still a member of a batch");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ctx_free(c);
    delete c;
    return PF_OK;
}

// ----------------------------------------------------------- stepping ----
template <class R>
Mesh<R> pf::make_view(const pf_ctx* c, int cur, int old, int nxt) {
    Mesh<R> m;
    memset(&m, 0, sizeof m);
    m.nx = c->nx; m.ny = c->ny; m.nz = c->nz; m.pz = c->pz; m.plane = c->plane;
    m.c0 = (const R*)c->lvl[cur][0]; m.c1 = (const R*)c->lvl[cur][1]; m.c2 = (const R*)c->lvl[cur][2];
    m.o0 = (const R*)c->lvl[old][0]; m.o1 = (const R*)c->lvl[old][1]; m.o2 = (const R*)c->lvl[old][2];
    m.n0 = (R*)c->lvl[nxt][0]; m.n1 = (R*)c->lvl[nxt][1]; m.n2 = (R*)c->lvl[nxt][2];
    m.cl0 = (const lo_t*)c->lvlo[cur][0]; m.cl1 = (const lo_t*)c->lvlo[cur][1]; m.cl2 = (const lo_t*)c->lvlo[cur][2];
    m.ol0 = (const lo_t*)c->lvlo[old][0]; m.ol1 = (const lo_t*)c->lvlo[old][1]; m.ol2 = (const lo_t*)c->lvlo[old][2];
    m.nl0 = (lo_t*)c->lvlo[nxt][0]; m.nl1 = (lo_t*)c->lvlo[nxt][1]; m.nl2 = (lo_t*)c->lvlo[nxt][2];
    m.ids = c->ids;
    m.mat = (const R*)c->mat;
    m.fx = (const R*)c->fx; m.fy = (const R*)c->fy; m.fz = (const R*)c->fz;
    m.sx = (const R*)c->sx; m.sy = (const R*)c->sy; m.sz = (const R*)c->sz;
    const bool exact = c->precision == PF_FP64;
    m.fx0 = (R)(exact ? c->f0[0] : 1.0 / c->f0[0]);
    m.fy0 = (R)(exact ? c->f0[1] : 1.0 / c->f0[1]);
    m.fz0 = (R)(exact ? c->f0[2] : 1.0 / c->f0[2]);
    m.sx0 = (R)(exact ? c->s0[0] : 1.0 / c->s0[0]);
    m.sy0 = (R)(exact ? c->s0[1] : 1.0 / c->s0[1]);
    m.sz0 = (R)(exact ? c->s0[2] : 1.0 / c->s0[2]);
    m.dpx = (const R*)c->dpx; m.dpy = (const R*)c->dpy; m.dpz = (const R*)c->dpz;
    for (int e = 0; e < 12; e++) {
        m.mur[e] = (R)c->mur[e];
        m.mur_lo[e] = (R)(c->mur[e] - (double)m.mur[e]);
    }
    m.wave = c->wave;
    m.step_base = nullptr;
    m.step_off = 0;
    m.nsteps = c->wave_len;
    m.nmat = c->nmat;
    m.surf = c->surface;
    return m;
}

template Mesh<float> pf::make_view<float>(const pf_ctx*, int, int, int);
template Mesh<double> pf::make_view<double>(const pf_ctx*, int, int, int);

// Launch the sweep of one step of mesh view m on stream s.
template <class R, int MODE>
static int launch_sweep_t(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    if constexpr (MODE == MODE_FP64P) {  // built for the one-step tiled sweep only
        if (c->kernel != PF_KERNEL_TILED) return fail(PF_ERR_STATE, "reciprocal spacings need the tiled sweep");
        return launch_tiled<R, MODE>(c, m, s);
    } else {
        if (c->kernel == PF_KERNEL_TILED || c->kernel == PF_KERNEL_TILED2) return launch_tiled<R, MODE>(c, m, s);
        if (c->kernel == PF_KERNEL_STRIP) return launch_strip<R, MODE>(c, m, s);
        dim3 blk(64, 4, 1);
        dim3 grd((unsigned)((c->nz + 63) / 64), (unsigned)((c->ny + 3) / 4), (unsigned)c->nx);
        if (c->sponge) hipLaunchKernelGGL((k_sweep_naive<R, MODE, true>), grd, blk, 0, s, m);
        else hipLaunchKernelGGL((k_sweep_naive<R, MODE, false>), grd, blk, 0, s, m);
        HIPCHK(hipGetLastError());
        return PF_OK;
    }
}

template <class R>
static int launch_probe_t(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    if (c->probe[2] <= 0) return PF_OK;
    // a step index known on the host (no device counter): launch only on
    // the steps that record a frame (the kernel checks it again)
    if (!m.step_base && (m.step_off >= m.nsteps || (m.step_off + 1) % c->every != 0)) return PF_OK;
    // newest level after this step = the n* buffers of the view
    const int pts = c->probe[2] * c->probe[5];
    hipLaunchKernelGGL((k_probe<R>), dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s, (const R*)m.n2,
                       (const lo_t*)m.nl2, c->plane, c->pz, c->probe[0], c->probe[1], c->probe[2], c->probe[3],
                       c->probe[4], c->probe[5], c->ring, c->cap, c->d_frame, (int*)(c->d_frame + 1),
                       m.step_base, m.step_off, c->every, m.nsteps);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

// Whether the step of view m records a probe frame: always possibly when its
// step index is on the device (the kernel decides), else by the host index.
template <class R>
static bool probe_due(const pf_ctx* c, const Mesh<R>& m) {
    if (c->probe[2] <= 0) return false;
    return m.step_base || (m.step_off < m.nsteps && (m.step_off + 1) % c->every == 0);
}

// The faces of the level view m writes, and the step's probe frame: one
// launch with the row-wise face kernel, which takes the probe blocks too;
// the per-cell face kernel and a probe launch on meshes too small for it.
template <class R, int MODE>
static int launch_faces_probe_t(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    long long rb, eb;
    if (face_rows_grid(c->nx, c->ny, c->nz, rb, eb)) {
        FaceProbe pr = {};
        pr.every = 1;
        if (probe_due(c, m)) {
            const int* q = c->probe;
            pr = {c->ring, c->d_frame, (int*)(c->d_frame + 1), q[0], q[1], q[2], q[3], q[4], q[5], c->cap, c->every,
                  (q[2] * q[5] + 255) / 256};
        }
        if (rb + eb + pr.blocks > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many face cells");
        hipLaunchKernelGGL((k_faces_rows<R, MODE>), dim3((unsigned)(rb + eb + pr.blocks)), dim3(256), 0, s, m, (int)rb,
                           (int)eb, pr);
        HIPCHK(hipGetLastError());
        return PF_OK;
    }
    const long long n = face_total(c->nx, c->ny, c->nz);
    hipLaunchKernelGGL((k_faces<R, MODE>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, n);
    HIPCHK(hipGetLastError());
    return launch_probe_t<R>(c, m, s);
}

// The view a sweep of MODE reads: with MODE_FP64P the spacing tables and
// first spacings are the exact reciprocals.
template <class R, int MODE>
static Mesh<R> sweep_view(const pf_ctx* c, int cur, int old, int nxt) {
    Mesh<R> m = make_view<R>(c, cur, old, nxt);
    if constexpr (MODE == MODE_FP64P) {
        m.fx = (const R*)c->rfx; m.fy = (const R*)c->rfy; m.fz = (const R*)c->rfz;
        m.sx = (const R*)c->rsx; m.sy = (const R*)c->rsy; m.sz = (const R*)c->rsz;
        m.fx0 = (R)(1.0 / c->f0[0]); m.fy0 = (R)(1.0 / c->f0[1]); m.fz0 = (R)(1.0 / c->f0[2]);
        m.sx0 = (R)(1.0 / c->s0[0]); m.sy0 = (R)(1.0 / c->s0[1]); m.sz0 = (R)(1.0 / c->s0[2]);
    }
    return m;
}

template <class R, int MODE>
static int step_once_t(pf_ctx* c, hipStream_t s, const long long* base, long long off,
                       hipEvent_t* ev /* 2 or null */) {
    Mesh<R> m = sweep_view<R, MODE>(c, c->cur, c->old, c->old);
    m.step_base = base;
    m.step_off = off;
    int rc;
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    if ((rc = launch_sweep_t<R, MODE>(c, m, s))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    // the faces read no spacings: the FP64 face kernels serve MODE_FP64P
    if ((rc = launch_faces_probe_t<R, MODE == MODE_FP64P ? MODE_FP64 : MODE>(c, m, s))) return rc;
    const int t = c->cur;
    c->cur = c->old;
    c->old = t;
    return PF_OK;
}

// Tile shape of the one-step tiled sweep by measurement, when the caller
// leaves it open: each built shape sweeps the new context's own fields (all
// zero at creation, step 0) on its stream, and the fastest is kept.  All
// shapes give the same bits, so only the speed depends on the choice.  Only
// the sweep runs (no faces, probes or step counters), and the level buffers
// are cleared afterwards.  Every shape first runs once untimed (the first
// launch of a kernel and the first touch of the fields cost extra), then
// the shapes are timed in two rounds, forward and backward, three sweeps
// each, and each keeps its faster round.  The best shape varies with the
// mesh (z extent against 62- or 30-lane tile rows, blocks per CU against
// the CU count) more than a fixed rule follows.
template <class R, int MODE>
static int tune_shape_t(pf_ctx* c) {
    static const int f32[][2] = {{16, 64}, {8, 64}, {32, 32}, {16, 32}, {8, 32}};
    static const int f64[][2] = {{8, 32}, {16, 32}, {8, 64}};
    const bool d64 = MODE == MODE_FP64 || MODE == MODE_FP64P;
    const int (*cand)[2] = d64 ? f64 : f32;
    const int nc = d64 ? 3 : 5;
    if (c->wave_len < 1) return PF_OK;  // the sweep would return at once
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreateWithFlags(&a, PF_TIMING_EVENT) != hipSuccess ||
        hipEventCreateWithFlags(&b, PF_TIMING_EVENT) != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        (void)hipGetLastError();
        return fail(PF_ERR_HIP, "tile tuning: cannot create events");
    }
    const Mesh<R> m = sweep_view<R, MODE>(c, c->cur, c->old, c->old);
    const int j0 = c->tile_j, z0 = c->tile_z;
    int rc = PF_OK;
    float best[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < nc && !rc; s++) {  // untimed: one sweep of every shape
        c->tile_j = cand[s][0];
        c->tile_z = cand[s][1];
        rc = launch_sweep_t<R, MODE>(c, m, c->stream);
    }
    for (int round = 0; round < 2 && !rc; round++) {
        for (int q = 0; q < nc && !rc; q++) {
            const int s = round == 0 ? q : nc - 1 - q;
            c->tile_j = cand[s][0];
            c->tile_z = cand[s][1];
            float ms = 0.f;
            if (hipEventRecord(a, c->stream) != hipSuccess) { rc = fail(PF_ERR_HIP, "tile tuning: event"); break; }
            for (int r = 0; r < 3 && !rc; r++) rc = launch_sweep_t<R, MODE>(c, m, c->stream);
            if (rc) break;
            if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                rc = fail(PF_ERR_HIP, "tile tuning: event timing");
                break;
            }
            if (round == 0 || ms < best[s]) best[s] = ms;
        }
    }
    int pick = 0;
    for (int s = 1; s < nc; s++)
        if (best[s] < best[pick]) pick = s;
    c->tile_j = rc ? j0 : cand[pick][0];
    c->tile_z = rc ? z0 : cand[pick][1];
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    const size_t lb = level_bytes(c->cells, c->precision);
    for (int l = 0; l < c->nlevels; l++) HIPCHK(hipMemsetAsync(c->lvl[l][0], 0, lb, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

static int tune_shape(pf_ctx* c) {
    switch (c->precision) {
        case PF_FP64: return c->recip ? tune_shape_t<double, MODE_FP64P>(c) : tune_shape_t<double, MODE_FP64>(c);
        case PF_FP32: return tune_shape_t<float, MODE_FP32>(c);
        default: return tune_shape_t<float, MODE_FP32C>(c);
    }
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
static int launch_pair_k(const pf_ctx* c, const Mesh<R>& m, R* out2, hipStream_t s) {
    static const int per_cu = blocks_per_cu(k_sweep_fused<R, MODE, SPONGE, TJ, KZ>, TJ * KZ);
    const TileMap tm = fused_map(c, TJ, KZ, c->tile_x, (long long)c->cus * per_cu);
    const long long nb = (long long)tm.nkt * tm.njt * tm.nxc;
    if (nb > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many tiles");
    hipLaunchKernelGGL((k_sweep_fused<R, MODE, SPONGE, TJ, KZ>), dim3((unsigned)nb), dim3(64, TJ * KZ / 64), 0, s, m,
                       out2, tm);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

template <class R, int MODE, bool SPONGE>
static int launch_pair_s(const pf_ctx* c, const Mesh<R>& m, R* out2, hipStream_t s) {
    const int tj = c->pair_j, kz = c->pair_z;
    if (kz == 64 && tj == 16) return launch_pair_k<R, MODE, SPONGE, 16, 64>(c, m, out2, s);
    if (kz == 32 && tj == 32) return launch_pair_k<R, MODE, SPONGE, 32, 32>(c, m, out2, s);
    return fail(PF_ERR_ARG, "two-step sweep: tile_j %d with tile_z %d is not a built shape", tj, kz);
}

// Steps tt, tt+1: the two-step pass writes level n+1 (view m1: n, n-1 ->
// a1) and the non-band cells of level n+2 (a2); then the level n+1 faces
// and probe frame, the band of level n+2 (view m2: n+1, n -> a2), its
// faces and probe frame.  a1, a2 are the two levels not in use.
template <class R, int MODE>
static int step_pair_t(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev) {
    int a1 = -1, a2 = -1;
    for (int l = 0; l < c->nlevels; l++)
        if (l != c->cur && l != c->old) (a1 < 0 ? a1 : a2) = l;
    if (a2 < 0) return fail(PF_ERR_STATE, "two-step sweep without its level buffers");
    Mesh<R> m1 = make_view<R>(c, c->cur, c->old, a1);
    m1.step_base = base;
    m1.step_off = off;
    Mesh<R> m2 = make_view<R>(c, a1, c->cur, a2);
    m2.step_base = base;
    m2.step_off = off + 1;
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    int rc = c->sponge ? launch_pair_s<R, MODE, true>(c, m1, m2.n0, s) : launch_pair_s<R, MODE, false>(c, m1, m2.n0, s);
    if (rc) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    if ((rc = launch_faces_probe_t<R, MODE>(c, m1, s))) return rc;
    const BandMap bm = band_map(c);
    const long long nb = band_blocks(bm);
    if (nb > 0x7fffffffLL) return fail(PF_ERR_ARG, "too many tiles");
    if (c->sponge)
        hipLaunchKernelGGL((k_band<R, MODE, true>), dim3((unsigned)nb), dim3(64, 4), 0, s, m2, bm.j, bm.i);
    else
        hipLaunchKernelGGL((k_band<R, MODE, false>), dim3((unsigned)nb), dim3(64, 4), 0, s, m2, bm.j, bm.i);
    HIPCHK(hipGetLastError());
    if ((rc = launch_faces_probe_t<R, MODE>(c, m2, s))) return rc;
    c->old = a1;
    c->cur = a2;
    return PF_OK;
}

int pf::step_pair(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev) {
    switch (c->precision) {
        case PF_FP32: return step_pair_t<float, MODE_FP32>(c, s, base, off, ev);
        case PF_FP32C: return step_pair_t<float, MODE_FP32C>(c, s, base, off, ev);
        default: return fail(PF_ERR_ARG, "the two-step sweep is built for PF_FP32 and PF_FP32C");
    }
}

int pf::step_once(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev) {
    switch (c->precision) {
        case PF_FP64:
            return c->recip ? step_once_t<double, MODE_FP64P>(c, s, base, off, ev)
                            : step_once_t<double, MODE_FP64>(c, s, base, off, ev);
        case PF_FP32: return step_once_t<float, MODE_FP32>(c, s, base, off, ev);
        default: return step_once_t<float, MODE_FP32C>(c, s, base, off, ev);
    }
}

extern "C" int pf_step(pf_ctx* c, int64_t nsteps) {
    if (!c) return fail(PF_ERR_ARG, "ctx is NULL");
    if (c->batch) return fail(PF_ERR_STATE, "context is stepped by its batch");
    if (nsteps < 0) return fail(PF_ERR_ARG, "nsteps < 0");
    if (c->step + nsteps > c->wave_len)
        return fail(PF_ERR_RANGE, "step %lld + %lld beyond the source table (%lld)", (long long)c->step,
                    (long long)nsteps, (long long)c->wave_len);
    int64_t s = 0;
    if (c->kernel == PF_KERNEL_TILED2) {
        for (; s + 2 <= nsteps; s += 2) {
            int rc = step_pair(c, c->stream, nullptr, c->step, nullptr);
            if (rc) return rc;
            c->step += 2;
        }
    }
    for (; s < nsteps; s++) {
        int rc = step_once(c, c->stream, nullptr, c->step, nullptr);
        if (rc) return rc;
        c->step++;
    }
    return PF_OK;
}

extern "C" int pf_sync(pf_ctx* c) {
    if (!c) return fail(PF_ERR_ARG, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

extern "C" int pf_steps_done(pf_ctx* c, int64_t* out) {
    if (!c || !out) return fail(PF_ERR_ARG, "NULL argument");
    *out = c->step;
    return PF_OK;
}

extern "C" int pf_kernel_info(pf_ctx* c, int32_t* kernel, int32_t* tile_j, int32_t* tile_z) {
    if (!c || !kernel || !tile_j || !tile_z) return fail(PF_ERR_ARG, "NULL argument");
    *kernel = c->kernel;
    const bool pair = c->kernel == PF_KERNEL_TILED2, naive = c->kernel == PF_KERNEL_NAIVE;
    *tile_j = naive ? 0 : (pair ? c->pair_j : c->tile_j);
    *tile_z = naive ? 0 : (pair ? c->pair_z : (c->kernel == PF_KERNEL_STRIP ? 64 : c->tile_z));
    return PF_OK;
}

// ---------------------------------------------------- field transfer ----
static void extents(const pf_ctx* c, int q, int& ex, int& ey, int& ez) {
    ex = c->nx - (q == 0);
    ey = c->ny - (q == 1);
    ez = c->nz - (q == 2);
}

static int ensure_scratch(pf_ctx* c, size_t bytes) {
    if (c->scratch_bytes >= bytes) return PF_OK;
    if (c->scratch) HIPCHK(hipFree(c->scratch));
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    HIPCHK(hipMalloc(&c->scratch, bytes));
    c->scratch_bytes = bytes;
    return PF_OK;
}

// ---------------------------------------------------- snapshots ----
extern "C" int pf_snapshot_open(pf_ctx* c, int32_t slots) {
    if (!c) return fail(PF_ERR_ARG, "ctx is NULL");
    if (slots < 1 || slots > PF_SNAP_MAX_SLOTS)
        return fail(PF_ERR_ARG, "slots must be 1..%d (got %d)", PF_SNAP_MAX_SLOTS, slots);
    if (c->snap_slots) return fail(PF_ERR_STATE, "snapshots are already open");
    long long frame = 0;
    for (int q = 0; q < 3; q++) {
        int ex, ey, ez;
        extents(c, q, ex, ey, ez);
        frame += (long long)ex * ey * ez;
    }
    const size_t bytes = (size_t)frame * sizeof(double) * (size_t)slots;
    HIPCHK(hipMalloc((void**)&c->snap_dev, bytes));
    HIPCHK(hipHostMalloc((void**)&c->snap_host, bytes, hipHostMallocDefault));
    HIPCHK(hipStreamCreateWithFlags(&c->snap_stream, hipStreamNonBlocking));
    for (int i = 0; i < slots; i++) {
        HIPCHK(hipEventCreateWithFlags(&c->snap_packed[i], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->snap_copied[i], hipEventDisableTiming));
        c->snap_state[i] = 0;
    }
    c->snap_frame = frame;
    c->snap_slots = slots;
    return PF_OK;
}

extern "C" int pf_snapshot(pf_ctx* c, int32_t slot) {
    if (!c) return fail(PF_ERR_ARG, "ctx is NULL");
    if (c->batch) return fail(PF_ERR_STATE, "context is stepped by its batch");
    if (slot < 0 || slot >= c->snap_slots) return fail(PF_ERR_ARG, "no snapshot slot %d", slot);
    if (c->snap_state[slot] != 0) return fail(PF_ERR_STATE, "snapshot slot %d is in use", slot);
    double* dev = c->snap_dev + (size_t)slot * (size_t)c->snap_frame;
    long long off = 0;
    for (int q = 0; q < 3; q++) {
        int ex, ey, ez;
        extents(c, q, ex, ey, ez);
        const long long n = (long long)ex * ey * ez;
        const unsigned g = (unsigned)((n + 255) / 256);
        if (c->esz == 8)
            hipLaunchKernelGGL((k_pack<double>), dim3(g), dim3(256), 0, c->stream, (const double*)c->lvl[c->cur][q],
                               (const lo_t*)nullptr, c->plane, c->pz, ex, ey, ez, dev + off, (const double*)nullptr);
        else
            hipLaunchKernelGGL((k_pack<float>), dim3(g), dim3(256), 0, c->stream, (const float*)c->lvl[c->cur][q],
                               (const lo_t*)c->lvlo[c->cur][q], c->plane, c->pz, ex, ey, ez, dev + off,
                               (const double*)nullptr);
        HIPCHK(hipGetLastError());
        off += n;
    }
    // the copy waits for the packing only; later steps on the context's
    // stream run while it is in flight (they do not touch the staging slot)
    HIPCHK(hipEventRecord(c->snap_packed[slot], c->stream));
    HIPCHK(hipStreamWaitEvent(c->snap_stream, c->snap_packed[slot], 0));
    HIPCHK(hipMemcpyAsync(c->snap_host + (size_t)slot * (size_t)c->snap_frame, dev,
                          (size_t)c->snap_frame * sizeof(double), hipMemcpyDeviceToHost, c->snap_stream));
    HIPCHK(hipEventRecord(c->snap_copied[slot], c->snap_stream));
    c->snap_state[slot] = 1;
    c->snap_step[slot] = c->step;
    return PF_OK;
}

extern "C" int pf_snapshot_wait(pf_ctx* c, int32_t slot, double** host, int64_t* steps_done) {
    if (!c || !host || !steps_done) return fail(PF_ERR_ARG, "NULL argument");
    if (slot < 0 || slot >= c->snap_slots) return fail(PF_ERR_ARG, "no snapshot slot %d", slot);
    if (c->snap_state[slot] == 0) return fail(PF_ERR_STATE, "snapshot slot %d holds no snapshot", slot);
    if (c->snap_state[slot] == 1) {
        HIPCHK(hipEventSynchronize(c->snap_copied[slot]));
        c->snap_state[slot] = 2;
    }
    *host = c->snap_host + (size_t)slot * (size_t)c->snap_frame;
    *steps_done = c->snap_step[slot];
    return PF_OK;
}

extern "C" int pf_snapshot_release(pf_ctx* c, int32_t slot) {
    if (!c) return fail(PF_ERR_ARG, "ctx is NULL");
    if (slot < 0 || slot >= c->snap_slots) return fail(PF_ERR_ARG, "no snapshot slot %d", slot);
    if (c->snap_state[slot] != 2) return fail(PF_ERR_STATE, "snapshot slot %d is not held (wait first)", slot);
    c->snap_state[slot] = 0;
    return PF_OK;
}

extern "C" int pf_get_ids(pf_ctx* c, uint8_t* host, int64_t count) {
    if (!c || !host) return fail(PF_ERR_ARG, "NULL argument");
    if (count != c->cells) return fail(PF_ERR_ARG, "the map has %lld nodes, got %lld", c->cells, (long long)count);
    if (c->batch) HIPCHK(hipStreamSynchronize(c->batch_stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, c->ids, (size_t)c->cells, hipMemcpyDeviceToHost));
    return PF_OK;
}

// Field `which` into host[0 .. count): packed to FP64 in the reference's
// shape into the device staging buffer `stage` (count doubles), then copied.
static int get_field_via(pf_ctx* c, int32_t which, double* host, int64_t count, double* stage) {
    const int q = which % 3;
    const int lvl = which < 3 ? c->cur : c->old;
    int ex, ey, ez;
    extents(c, q, ex, ey, ez);
    const long long n = (long long)ex * ey * ez;
    // the previous level of uz holds the source value on its wall line, as
    // the reference's u_old does after time_step
    const double* wsub = (which == PF_UZ_OLD && c->step > 0) ? c->wave + (c->step - 1) : nullptr;
    unsigned g = (unsigned)((n + 255) / 256);
    if (c->esz == 8)
        hipLaunchKernelGGL((k_pack<double>), dim3(g), dim3(256), 0, c->stream, (const double*)c->lvl[lvl][q],
                           (const lo_t*)nullptr, c->plane, c->pz, ex, ey, ez, stage, wsub);
    else
        hipLaunchKernelGGL((k_pack<float>), dim3(g), dim3(256), 0, c->stream, (const float*)c->lvl[lvl][q],
                           (const lo_t*)c->lvlo[lvl][q], c->plane, c->pz, ex, ey, ez, stage, wsub);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, stage, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)count;
    return PF_OK;
}

static int check_field(pf_ctx* c, int32_t which, const void* host, int64_t count, long long* n_out) {
    if (!c || !host) return fail(PF_ERR_ARG, "NULL argument");
    if (which < 0 || which > 5) return fail(PF_ERR_ARG, "bad field %d", which);
    int ex, ey, ez;
    extents(c, which % 3, ex, ey, ez);
    const long long n = (long long)ex * ey * ez;
    if (count != n) return fail(PF_ERR_ARG, "field %d has %lld elements, got %lld", which, n, (long long)count);
    *n_out = n;
    return PF_OK;
}

extern "C" int pf_get_field(pf_ctx* c, int32_t which, double* host, int64_t count) {
    long long n = 0;
    int rc = check_field(c, which, host, count, &n);
    if (rc) return rc;
    if (c->batch) HIPCHK(hipStreamSynchronize(c->batch_stream));
    if ((rc = ensure_scratch(c, (size_t)n * sizeof(double)))) return rc;
    return get_field_via(c, which, host, count, (double*)c->scratch);
}

// ------------------------------------------- allocation-free transfer ----
// One staging array per device for pf_get_field_noalloc, held for the life
// of the process.  Its name is a host-side stand-in: host code uses only the
// device address hipGetSymbolAddress gives for the current device, taken
// once per device when a context is created there (noalloc_scratch_init).
static __device__ __align__(256) double scratch[PF_NOALLOC_FIELD_MAX];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

int pf::noalloc_scratch_init(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(PF_ERR_ARG, "device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return PF_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the current device is dev
    g_scratch_ptr[dev] = (double*)p;
    return PF_OK;
}

extern "C" int pf_get_field_noalloc(pf_ctx* c, int32_t which, double* host, int64_t count) {
    long long n = 0;
    int rc = check_field(c, which, host, count, &n);
    if (rc) return rc;
    if (n > PF_NOALLOC_FIELD_MAX)
        return fail(PF_ERR_ARG, "field %d has %lld elements; pf_get_field_noalloc takes at most %lld (use pf_get_field)",
                    which, n, (long long)PF_NOALLOC_FIELD_MAX);
    double* stage = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        stage = (c->dev >= 0 && c->dev < MAX_DEVICES) ? g_scratch_ptr[c->dev] : nullptr;
    }
    if (!stage) return fail(PF_ERR_STATE, "no staging array on device %d", c->dev);
    if (c->batch) HIPCHK(hipStreamSynchronize(c->batch_stream));
    return get_field_via(c, which, host, count, stage);
}

extern "C" int pf_set_field(pf_ctx* c, int32_t which, const double* host, int64_t count) {
    if (!c || !host) return fail(PF_ERR_ARG, "NULL argument");
    if (which < 0 || which > 5) return fail(PF_ERR_ARG, "bad field %d", which);
    if (c->batch) return fail(PF_ERR_STATE, "context is stepped by its batch");
    const int q = which % 3;
    const int lvl = which < 3 ? c->cur : c->old;
    int ex, ey, ez;
    extents(c, q, ex, ey, ez);
    const long long n = (long long)ex * ey * ez;
    if (count != n) return fail(PF_ERR_ARG, "field %d has %lld elements, got %lld", which, n, (long long)count);
    int rc = ensure_scratch(c, (size_t)n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->scratch, host, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    unsigned g = (unsigned)((n + 255) / 256);
    if (c->esz == 8)
        hipLaunchKernelGGL((k_unpack<double>), dim3(g), dim3(256), 0, c->stream, (double*)c->lvl[lvl][q],
                           (lo_t*)nullptr, c->plane, c->pz, ex, ey, ez, (const double*)c->scratch);
    else
        hipLaunchKernelGGL((k_unpack<float>), dim3(g), dim3(256), 0, c->stream, (float*)c->lvl[lvl][q],
                           (lo_t*)c->lvlo[lvl][q], c->plane, c->pz, ex, ey, ez, (const double*)c->scratch);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return PF_OK;
}

// ------------------------------------------------------------- probes ----
extern "C" int pf_get_probes(pf_ctx* c, double* host, int64_t cap_frames, int64_t* nframes) {
    if (!c || !nframes) return fail(PF_ERR_ARG, "NULL argument");
    *nframes = 0;
    if (c->probe[2] <= 0) return PF_OK;
    hipStream_t s = c->batch ? c->batch_stream : c->stream;
    HIPCHK(hipStreamSynchronize(s));
    long long f = 0;
    HIPCHK(hipMemcpy(&f, c->d_frame, sizeof f, hipMemcpyDeviceToHost));
    const long long per = (long long)c->probe[2] * c->probe[5];
    const long long have = f < c->cap ? f : c->cap;
    const long long first = f - have;  // oldest frame still in the ring
    const long long take = have < cap_frames ? have : cap_frames;
    if (take > 0 && !host) return fail(PF_ERR_ARG, "host is NULL");
    for (long long t = 0; t < take; t++) {
        long long slot = (first + t) % c->cap;
        HIPCHK(hipMemcpy(host + t * per, c->ring + slot * per, (size_t)per * sizeof(double), hipMemcpyDeviceToHost));
    }
    *nframes = take;
    return PF_OK;
}

// ------------------------------------------------------------- energy ----
template <class R>
__global__ void k_widen(const R* __restrict__ hi, const lo_t* __restrict__ lo, double* __restrict__ out, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const R h = hi[t];
    double v = (double)h;
    if (lo) v += (double)lo_decode(h, lo[t]);
    out[t] = v;
}

extern "C" int pf_energy(pf_ctx* c, double* out) {
    if (!c || !out) return fail(PF_ERR_ARG, "NULL argument");
    hipStream_t s = c->batch ? c->batch_stream : c->stream;
    EnergyView ev;
    memset(&ev, 0, sizeof ev);
    const int newest = c->cur, prev = c->old;
    if (c->esz == 8) {
        ev.m = make_view<double>(c, newest, prev, prev);
    } else {
        int rc = ensure_scratch(c, (size_t)c->cells * sizeof(double) * 6);
        if (rc) return rc;
        double* w = (double*)c->scratch;
        unsigned g = (unsigned)((c->cells + 255) / 256);
        for (int l = 0; l < 2; l++)
            for (int q = 0; q < 3; q++) {
                int lv = l == 0 ? newest : prev;
                hipLaunchKernelGGL((k_widen<float>), dim3(g), dim3(256), 0, s, (const float*)c->lvl[lv][q],
                                   (const lo_t*)c->lvlo[lv][q], w + (size_t)(3 * l + q) * c->cells, c->cells);
            }
        HIPCHK(hipGetLastError());
        Mesh<double>& m = ev.m;
        memset(&m, 0, sizeof m);
        m.nx = c->nx; m.ny = c->ny; m.nz = c->nz; m.pz = c->pz; m.plane = c->plane;
        m.c0 = w; m.c1 = w + c->cells; m.c2 = w + 2 * c->cells;
        m.n0 = w + 3 * c->cells; m.n1 = w + 4 * c->cells; m.n2 = w + 5 * c->cells;
        m.ids = c->ids;
    }
    Mesh<double>& m = ev.m;
    m.mat = (const double*)c->mat64;
    m.fx = (const double*)c->fx64; m.fy = (const double*)c->fy64; m.fz = (const double*)c->fz64;
    m.sx = (const double*)c->sx64; m.sy = (const double*)c->sy64; m.sz = (const double*)c->sz64;
    m.fx0 = c->f0[0]; m.fy0 = c->f0[1]; m.fz0 = c->f0[2];
    m.sx0 = c->s0[0]; m.sy0 = c->s0[1]; m.sz0 = c->s0[2];
    ev.rho = (const double*)c->rho64;
    ev.dt = c->dt;
    const long long done = c->step;
    ev.subst_prev = done > 0 ? 1 : 0;
    if (done > 0) HIPCHK(hipMemcpy(&ev.w_prev, c->wave + (done - 1), sizeof(double), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_energy_partial, dim3(c->e_blocks), dim3(256), 0, s, ev, c->e_partials);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, (const double*)c->e_partials, c->e_blocks, c->e_out);
    HIPCHK(hipGetLastError());
    double h[2];
    HIPCHK(hipMemcpyAsync(h, c->e_out, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *out = 0.5 * c->dV * (h[0] + h[1]);
    return PF_OK;
}

// ------------------------------------------------------------ profile ----
extern "C" int pf_profile(pf_ctx* c, int64_t nsteps, double* sweep_ms, double* face_ms) {
    if (!c || !sweep_ms || !face_ms) return fail(PF_ERR_ARG, "NULL argument");
    if (c->batch) return fail(PF_ERR_STATE, "context is stepped by its batch");
    if (nsteps < 1) return fail(PF_ERR_ARG, "nsteps < 1");
    if (c->step + nsteps > c->wave_len) return fail(PF_ERR_RANGE, "beyond the source table");
    // Launch groups: a pair's two-step pass with its band, faces and probes,
    // or a single step.  Events bracket the sweep of every stride-th group
    // of each kind (the first included; stride 4, or more so that at most
    // PF_PROFILE_BRACKETS groups of a kind are bracketed), and one event
    // follows the last group: a recorded event is a marker between two
    // launches that holds the next launch back by a few microseconds, so
    // bracketing every group would slow the steps it times.  Sweep time =
    // the mean bracketed sweep of each kind times the groups of that kind;
    // the rest (faces, probes, the band of a pair, launch gaps) = the span
    // from the first event to the last less the sweeps.  The events are the
    // context's own, created by the first call.
    const int64_t pairs = c->kernel == PF_KERNEL_TILED2 ? nsteps / 2 : 0;
    const int64_t singles = nsteps - 2 * pairs;
    auto stride_of = [](int64_t groups) {
        const int64_t s = (groups + PF_PROFILE_BRACKETS - 1) / PF_PROFILE_BRACKETS;
        return s > 4 ? s : (int64_t)4;
    };
    const int64_t pstride = stride_of(pairs), sstride = stride_of(singles);
    hipEvent_t* ev = c->prof_ev;
    if (!ev[PF_PROFILE_EVENTS - 1]) {
        for (auto& e : c->prof_ev) {
            if (e) continue;
            const hipError_t err = hipEventCreateWithFlags(&e, PF_TIMING_EVENT);
            if (err != hipSuccess) {
                e = nullptr;
                return fail(PF_ERR_HIP, "hipEventCreateWithFlags: %s", hipGetErrorString(err));
            }
        }
    }
    int rc = PF_OK;
    size_t used = 0;
    for (int64_t g = 0; g < pairs && !rc; g++) {
        hipEvent_t* e = g % pstride == 0 ? &ev[used] : nullptr;
        if (e) used += 2;
        rc = step_pair(c, c->stream, nullptr, c->step, e);
        if (!rc) c->step += 2;
    }
    const size_t pair_ev = used;
    for (int64_t g = 0; g < singles && !rc; g++) {
        hipEvent_t* e = g % sstride == 0 ? &ev[used] : nullptr;
        if (e) used += 2;
        rc = step_once(c, c->stream, nullptr, c->step, e);
        if (!rc) c->step++;
    }
    if (!rc) {
        HIPCHK(hipEventRecord(ev[used], c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double sum[2] = {0, 0};
        for (size_t g = 0; g < used; g += 2) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ev[g], ev[g + 1]));
            sum[g < pair_ev ? 0 : 1] += t;
        }
        const double np = (double)(pair_ev / 2), ns = (double)((used - pair_ev) / 2);
        const double sw = (np > 0 ? sum[0] / np * (double)pairs : 0.0) + (ns > 0 ? sum[1] / ns * (double)singles : 0.0);
        float all = 0;
        HIPCHK(hipEventElapsedTime(&all, ev[0], ev[used]));
        *sweep_ms = sw / (double)nsteps;
        *face_ms = ((double)all - sw) / (double)nsteps;
    }
    return rc;
}
