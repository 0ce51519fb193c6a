// Host-side context of libphonofdtd, shared by the API translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/phonofdtd.h"
#include "pf_kernels.hpp"

// Flags of the events the library times launches with: no system-scope
// fence when an event is recorded (a recorded event would otherwise write
// back and invalidate the caches between the launches it brackets, which
// slows the launches after it; the elapsed times are read only after a
// stream synchronization, which makes the event visible).
#define PF_TIMING_EVENT hipEventDisableSystemFence

// Material rows the tiled sweep stages in LDS.
#define PF_TILED_MAX_MATERIALS 16
// Snapshot slots a context can hold.
#define PF_SNAP_MAX_SLOTS 16
// Offset of time level l in its allocation: l * PF_LEVEL_STAGGER bytes (an
// odd multiple of 256).
#define PF_LEVEL_STAGGER 4352
// Launch groups of each kind (pairs, single steps) that one pf_profile call
// brackets with events at most.
#define PF_PROFILE_BRACKETS 8
#define PF_PROFILE_EVENTS (4 * PF_PROFILE_BRACKETS + 1)

// One mesh on one device.  Time levels of the displacement components are
// kept in level buffers: `cur` indexes the current level (n) and `old` the
// previous one (n-1); a single step overwrites `old` with level n+1 and
// swaps the two.  Two-step sweeps use two more level buffers (allocated on
// first use).  Each level is one allocation holding its arrays back to back
// (ux, uy, uz, then the 16-bit lo parts in FP32C), so one buffer descriptor
// covers a whole level; level l starts l * PF_LEVEL_STAGGER bytes into its
// allocation (lvlbase[l] is the allocation, lvl[l][0] the level).
struct pf_ctx {
    int nx = 0, ny = 0, nz = 0, pz = 0;
    long long plane = 0, cells = 0;  // ny*pz, nx*ny*pz
    int precision = 0, kernel = 0, tile_j = 4, tile_x = 0, tile_z = 64;
    int pair_j = 0, pair_z = 0;  // tile of the two-step sweep (PF_KERNEL_TILED2)
    bool shape_open = false;     // tile shape left to measurement at creation
    size_t esz = 8;
    bool sponge = false;
    int surface = 0;  // pf_desc.surface
    double dt = 0;
    int nmat = 0;
    int dev = 0;
    int cus = 0;  // compute units of the device (0: unknown)
    hipStream_t stream = nullptr;
    void* lvlbase[4] = {};
    void* lvl[4][3] = {};
    void* lvlo[4][3] = {};
    int nlevels = 0;
    int cur = 0, old = 1;
    uint8_t* ids = nullptr;
    void *mat = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr, *sx = nullptr, *sy = nullptr,
         *sz = nullptr;
    void *dpx = nullptr, *dpy = nullptr, *dpz = nullptr;
    // FP64 with every spacing a power of two: the reciprocal tables of the
    // one-step tiled sweep's products (MODE_FP64P); fx.. keep the spacings
    // for the other kernels (batch, energy).
    bool recip = false;
    void *rfx = nullptr, *rfy = nullptr, *rfz = nullptr, *rsx = nullptr, *rsy = nullptr, *rsz = nullptr;
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
    long long* d_frame = nullptr;  // [0] frames recorded; [1], as an int: blocks done in the probe launch
    double* e_out = nullptr;
    double* e_partials = nullptr;
    int e_blocks = 0;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    pf_batch* batch = nullptr;  // set while a batch owns the stepping
    hipStream_t batch_stream = nullptr;
    // asynchronous snapshots: per slot a device staging frame and a pinned
    // host frame (ux, uy, uz in the reference's shapes, FP64), the copy on
    // its own stream
    int snap_slots = 0;
    long long snap_frame = 0;  // doubles per frame
    double* snap_dev = nullptr;
    double* snap_host = nullptr;
    hipStream_t snap_stream = nullptr;
    hipEvent_t snap_packed[PF_SNAP_MAX_SLOTS] = {};
    hipEvent_t snap_copied[PF_SNAP_MAX_SLOTS] = {};
    int snap_state[PF_SNAP_MAX_SLOTS] = {};      // 0 free, 1 in flight, 2 held by the caller
    long long snap_step[PF_SNAP_MAX_SLOTS] = {};
    // pf_profile's timing events, created by its first call and kept, so
    // that later calls time the steps and not the event set-up
    hipEvent_t prof_ev[PF_PROFILE_EVENTS] = {};
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
// View of a step from level `cur` (n) and `old` (n-1) into level `out`.
template <class R>
Mesh<R> make_view(const pf_ctx* c, int cur, int old, int out);
// Issue one step (sweep, faces, probe) on stream s; tt = *base + off (or
// off when base is null).  ev: 2 events recorded before and after the
// sweep, or null.
int step_once(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev);
// Issue steps tt and tt+1 with the two-step sweep (PF_KERNEL_TILED2); ev as
// for step_once (around the two-step pass).
int step_pair(pf_ctx* c, hipStream_t s, const long long* base, long long off, hipEvent_t* ev);
// Device address of the allocation-free staging array on device dev, taken
// once per device (called at context creation, with dev current).
int noalloc_scratch_init(int dev);

}  // namespace pf

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            return pf::fail(PF_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                            __LINE__);                                                              \
    } while (0)
