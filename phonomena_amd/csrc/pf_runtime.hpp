// Host-side context of libphonofdtd, shared by the API translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/phonofdtd.h"
#include "pf_kernels.hpp"

// Material rows the tiled sweep stages in LDS.
#define PF_TILED_MAX_MATERIALS 16

// One mesh on one device.  Two levels of each displacement component are
// kept (ping-pong); `cur` indexes the current level.  Each level is one
// allocation holding its arrays back to back (ux, uy, uz, then the lo parts
// in FP32C), so one buffer descriptor covers a whole level.
struct pf_ctx {
    int nx = 0, ny = 0, nz = 0, pz = 0;
    long long plane = 0, cells = 0;  // ny*pz, nx*ny*pz
    int precision = 0, kernel = 0, tile_j = 4, tile_x = 0, tile_z = 64;
    size_t esz = 8;
    bool sponge = false;
    double dt = 0;
    int nmat = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    void* lvlbase[2] = {};
    void* lvl[2][3] = {};
    void* lvlo[2][3] = {};
    int cur = 0;
    uint8_t* ids = nullptr;
    void *mat = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr, *sx = nullptr, *sy = nullptr,
         *sz = nullptr;
    void *dpx = nullptr, *dpy = nullptr, *dpz = nullptr;
    void *mat64 = nullptr, *rho64 = nullptr, *fx64 = nullptr, *fy64 = nullptr, *fz64 = nullptr,
         *sx64 = nullptr, *sy64 = nullptr, *sz64 = nullptr;
    double f0[3] = {}, s0[3] = {};
    double dV = 0;
    double mur[12] = {};
    double* wave = nullptr;
    long long wave_len = 0;
    long long* d_step = nullptr;  // device step counter (graph replays)
    long long step = 0;           // steps issued
    int probe[6] = {};
    int every = 0, cap = 0;
    double* ring = nullptr;
    long long* d_frame = nullptr;
    double* e_out = nullptr;
    double* e_partials = nullptr;
    int e_blocks = 0;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    pf_batch* batch = nullptr;  // set while a batch owns the stepping
    hipStream_t batch_stream = nullptr;
};

namespace pf {

// Bytes of one time level: three components of esz bytes, plus three
// 16-bit lo parts in PF_FP32C.
inline unsigned long long level_bytes(long long cells, int precision) {
    const unsigned long long c = (unsigned long long)cells;
    if (precision == PF_FP64) return c * 3 * 8;
    return c * 3 * 4 + (precision == PF_FP32C ? c * 3 * sizeof(lo_t) : 0);
}

int fail(int code, const char* fmt, ...);
int ctx_init(pf_ctx* c, const pf_desc* d);
void ctx_free(pf_ctx* c);
template <class R>
Mesh<R> make_view(const pf_ctx* c, int cur);
// Issue one step (sweep, faces, probe) on stream s; tt = *base + off (or
// off when base is null).  ev: 3 events recorded around the launches, or null.
int step_once(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev);

}  // namespace pf

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            return pf::fail(PF_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                            __LINE__);                                                              \
    } while (0)
