// Batched stepping: many contexts of one precision stepped together.
//
// Per step the batch issues three launches whatever its size:
//   1. one persistent sweep over a combined tile work-list of all members
//      (the same per-tile code as the single-mesh tiled sweep, so a member's
//      fields are bit-identical to stepping it alone);
//   2. one launch over all members' boundary-face cells;
//   3. one launch recording every member's probe frame, whose last block to
//      finish advances every member's device step counter.
// The step index of a member lives on the device, and which of its two
// levels is current follows from it, so K steps are captured once in a HIP
// graph and replayed.  A member whose step counter reaches its source-table
// length (its step budget) stops: its tiles, faces and probes return at
// once and its counter no longer moves.
//
// Work-list: the tiles of every member, member after member in tile order,
// cut into 8 contiguous runs of equal plane count, one per XCD (the device
// hands out workgroups to XCDs round-robin, so block b runs on XCD b % 8).
// Each block takes the next tile of its XCD's run from an atomic counter,
// so neighbouring tiles, which share halo rows, are swept on the same L2.
// The counters of a step are reset by the previous step's launch.
//
// Tile shapes (rows x lanes and planes per chunk) are the batch's; the
// autotuner times candidate shapes on real steps (every shape gives the
// same bits) and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "pf_runtime.hpp"
#include "pf_tiled.hpp"

using namespace pf;

namespace pf {

constexpr int NXCD = 8;

template <class R>
struct BMember {
    Mesh<R> view[2];        // mesh views with level 0 / level 1 current
    TileMap tm;
    long long* step;        // the member's device step counter
    long long t0;           // its step count when the views were taken
    int cur0;               // and its current level then
    int probe[6];
    int every, cap;
    double* ring;
    long long* frame;
    long long faces;        // boundary-face cells per step
    long long budget;       // source-table length
};

// The member table is read through the constant address space: its loads
// are scalar (one per block, into scalar registers), as a kernel argument's
// are.  Through a plain pointer every use of a mesh field inside the march
// became a vector load with a full wait.
template <class R>
using CMember = const __attribute__((address_space(4))) BMember<R>;

struct WorkItem {
    int member, tile;
};
typedef const __attribute__((address_space(4))) WorkItem CItem;

// A copy of *p read word by word through the constant address space
// (block-uniform p), i.e. with scalar loads.
template <class T>
PF_DEV T load_const(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized struct");
    typedef const __attribute__((address_space(4))) unsigned int CWord;
    const CWord* w = (const CWord*)p;
    T v;
    unsigned int* d = (unsigned int*)&v;
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; k++) d[k] = w[k];
    return v;
}

// Mesh view of member i at its current step (the step counter is written
// only between launches, so one read per block is current for the launch).
template <class R>
PF_DEV Mesh<R> current_view(const BMember<R>* mem, int i) {
    CMember<R>* cm = (CMember<R>*)mem;
    const long long tt = *cm[i].step;
    const int p = cm[i].cur0 ^ (int)((tt - cm[i].t0) & 1);
    return load_const(&mem[i].view[p]);
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
__global__ __launch_bounds__(TJ * KZ, (MODE == MODE_FP32 ? 4 : 2)) void k_batch_sweep(
    const BMember<R>* __restrict__ mem, const WorkItem* __restrict__ items, const int* __restrict__ qbeg,
    int* __restrict__ ctr, const long long* __restrict__ bstep) {
    __shared__ TiledLds<R, TJ, KZ> L;
    __shared__ int s_item;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int x = (int)(blockIdx.x % NXCD);
    const int par = (int)(*bstep & 1);
    if (blockIdx.x == 0 && tid < NXCD) ctr[(par ^ 1) * NXCD + tid] = 0;  // next step's counters
    int* c = ctr + par * NXCD + x;
    const int end = qbeg[x + 1];
    for (;;) {
        __syncthreads();  // the previous tile's LDS reads and s_item are done
        if (tid == 0) s_item = qbeg[x] + atomicAdd(c, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= end) break;  // uniform: every wave of the block leaves
        CItem* ci = (CItem*)items;
        const int mi = ci[it].member, tile = ci[it].tile;
        CMember<R>* cm = (CMember<R>*)mem;
        const TileMap tm = {cm[mi].tm.nkt, cm[mi].tm.njt, cm[mi].tm.nxc, cm[mi].tm.xchunk};
        const Mesh<R> m = current_view(mem, mi);
        tiled_tile<R, MODE, SPONGE, TJ, KZ>(m, tm, tile, L);
    }
}

// Face block g of a batched step: its member and its index among that
// member's face blocks (one table load instead of a search).
struct FaceBlock {
    int member, block;
};

// Face cells of all members, BATCH_FACE_CELLS per block (four per thread).
constexpr int BATCH_FACE_CELLS = 1024;
template <class R, int MODE>
__global__ __launch_bounds__(256) void k_batch_faces(const BMember<R>* __restrict__ mem,
                                                     const FaceBlock* __restrict__ fmap) {
    const FaceBlock fb = load_const(&fmap[blockIdx.x]);
    CMember<R>* cm = (CMember<R>*)mem;
    const Mesh<R> m = current_view(mem, fb.member);
    faces_run<R, MODE, BATCH_FACE_CELLS / 256>(m, (long long)fb.block * BATCH_FACE_CELLS, cm[fb.member].faces);
}

// Block i records member i's probe frame.  Every block then counts itself
// done; the last one (which sees every other block's count, and so every
// block's reads of its step counter, before it) advances all members' step
// counters, the batch's step parity, and resets the count for the next
// step.  No block waits for another.
template <class R>
__global__ __launch_bounds__(256) void k_batch_probe(const BMember<R>* __restrict__ mem,
                                                     long long* const* __restrict__ steps,
                                                     const long long* __restrict__ budget, int nmem,
                                                     long long* __restrict__ bstep, int* __restrict__ done) {
    __shared__ int s_last;
    CMember<R>* cm = (CMember<R>*)mem;
    const int i = (int)blockIdx.x;
    if (cm[i].probe[2] > 0) {
        const Mesh<R> m = current_view(mem, i);
        probe_record<R>(m.n2, m.nl2, m.plane, m.pz, cm[i].probe[0], cm[i].probe[1], cm[i].probe[2],
                        cm[i].probe[3], cm[i].probe[4], cm[i].probe[5], cm[i].ring, cm[i].cap, cm[i].frame,
                        *cm[i].step, cm[i].every, cm[i].budget);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(done, 1) == nmem - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (int k = threadIdx.x; k < nmem; k += blockDim.x) {
        const long long v = *steps[k] + 1;
        *steps[k] = v < budget[k] ? v : budget[k];
    }
    if (threadIdx.x == 0) {
        *bstep += 1;
        *done = 0;
    }
}

}  // namespace pf

struct pf_batch {
    int precision = 0;
    int sponge = -1;  // set by the first member
    std::vector<pf_ctx*> members;
    hipStream_t stream = nullptr;
    int dev = 0, cus = 0;
    int tile_j = 0, tile_z = 0, tile_x = 0;
    bool use_graph = true;
    // device tables
    void* d_mem = nullptr;
    size_t mem_bytes = 0;
    WorkItem* d_items = nullptr;
    size_t items_cap = 0;
    int* d_qbeg = nullptr;
    int* d_ctr = nullptr;
    long long* d_bstep = nullptr;
    FaceBlock* d_fmap = nullptr;
    size_t fmap_cap = 0;
    long long** d_steps = nullptr;
    long long* d_budget = nullptr;
    size_t nmem_cap = 0;
    long long face_blocks = 0;
    int grid = 0;
    bool dirty = true;
    // captured step loop
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
};

namespace {

int free_graph(pf_batch* b) {
    if (b->exec) HIPCHK(hipGraphExecDestroy(b->exec));
    if (b->graph) HIPCHK(hipGraphDestroy(b->graph));
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
    return PF_OK;
}

void free_tables(pf_batch* b) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(b->d_mem); F(b->d_items); F(b->d_qbeg); F(b->d_ctr); F(b->d_bstep); F(b->d_fmap);
    F(b->d_steps); F(b->d_budget);
    b->d_mem = nullptr; b->d_items = nullptr; b->d_qbeg = nullptr; b->d_ctr = nullptr; b->d_bstep = nullptr;
    b->d_fmap = nullptr; b->d_steps = nullptr; b->d_budget = nullptr;
    b->mem_bytes = b->items_cap = b->nmem_cap = b->fmap_cap = 0;
}

// Default shapes of a batch (the autotuner refines them).
void default_shape(const pf_batch* b, int& tj, int& tz, int& tx) {
    tj = b->tile_j;
    tz = b->tile_z;
    tx = b->tile_x;
    if (tz == 0) tz = b->precision == PF_FP64 ? 32 : 64;
    if (tj == 0) tj = b->precision == PF_FP64 ? 8 : (tz == 64 ? 16 : 32);
    if (tx == 0) tx = 32;
}

bool shape_built(int tj, int tz) {
    return (tz == 64 && (tj == 8 || tj == 16)) || (tz == 32 && (tj == 8 || tj == 16 || tj == 32));
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
int sweep_grid(pf_batch* b) {
    int per = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_batch_sweep<R, MODE, SPONGE, TJ, KZ>,
                                                        TJ * KZ, 0));
    if (per < 1) per = 1;
    b->grid = per * b->cus;
    b->grid = (b->grid + NXCD - 1) / NXCD * NXCD;
    return PF_OK;
}

template <class R, int MODE, bool SPONGE>
int sweep_grid_s(pf_batch* b, int tj, int tz) {
    if (tz == 64 && tj == 16) return sweep_grid<R, MODE, SPONGE, 16, 64>(b);
    if (tz == 64 && tj == 8) return sweep_grid<R, MODE, SPONGE, 8, 64>(b);
    if (tz == 32 && tj == 32) return sweep_grid<R, MODE, SPONGE, 32, 32>(b);
    if (tz == 32 && tj == 16) return sweep_grid<R, MODE, SPONGE, 16, 32>(b);
    return sweep_grid<R, MODE, SPONGE, 8, 32>(b);
}

template <class R, int MODE>
int build_t(pf_batch* b) {
    int tj, tz, tx;
    default_shape(b, tj, tz, tx);
    const int nmem = (int)b->members.size();
    std::vector<BMember<R>> mem((size_t)nmem);
    std::vector<WorkItem> items;
    std::vector<long long> cost;  // planes per work item
    std::vector<long long> fblk((size_t)nmem + 1, 0);
    std::vector<long long*> steps((size_t)nmem);
    std::vector<long long> budget((size_t)nmem);
    for (int i = 0; i < nmem; i++) {
        pf_ctx* c = b->members[(size_t)i];
        BMember<R>& e = mem[(size_t)i];
        memset(&e, 0, sizeof e);
        // view[0]: the level current now; view[1]: after an odd count of steps
        e.view[0] = make_view<R>(c, c->cur, c->old, c->old);
        e.view[1] = make_view<R>(c, c->old, c->cur, c->cur);
        for (int p = 0; p < 2; p++) {
            e.view[p].step_base = c->d_step;
            e.view[p].step_off = 0;
        }
        // planes per chunk: about tx, evened out over the mesh
        const int nch = (c->nx + tx - 1) / tx;
        const int xchunk = std::min((c->nx + nch - 1) / nch, XT - 3);
        e.tm = tile_map(c, tj, tz, xchunk);
        e.step = c->d_step;
        e.t0 = c->step;
        e.cur0 = 0;
        for (int q = 0; q < 6; q++) e.probe[q] = c->probe[q];
        e.every = c->every;
        e.cap = c->cap;
        e.ring = c->ring;
        e.frame = c->d_frame;
        e.faces = face_total(c->nx, c->ny, c->nz);
        e.budget = c->wave_len;
        const long long nt = (long long)e.tm.nkt * e.tm.njt * e.tm.nxc;
        for (long long t = 0; t < nt; t++) {
            items.push_back({i, (int)t});
            const int xc = (int)(t / ((long long)e.tm.nkt * e.tm.njt));
            const int ib = xc * e.tm.xchunk;
            cost.push_back(std::min(e.tm.xchunk, c->nx - ib) + 2);
        }
        fblk[(size_t)i + 1] = fblk[(size_t)i] + (e.faces + BATCH_FACE_CELLS - 1) / BATCH_FACE_CELLS;
        steps[(size_t)i] = c->d_step;
        budget[(size_t)i] = c->wave_len;
    }
    if (items.size() > 0x7fffffff) return fail(PF_ERR_ARG, "batch work-list too long");
    if (fblk[(size_t)nmem] > 0x7fffffffLL) return fail(PF_ERR_ARG, "batch has too many face cells");
    std::vector<FaceBlock> fmap;
    fmap.reserve((size_t)fblk[(size_t)nmem]);
    for (int i = 0; i < nmem; i++)
        for (long long g = fblk[(size_t)i]; g < fblk[(size_t)i + 1]; g++)
            fmap.push_back({i, (int)(g - fblk[(size_t)i])});
    // 8 contiguous runs of (nearly) equal plane count
    long long total = 0;
    for (long long v : cost) total += v;
    std::vector<int> qbeg(NXCD + 1, 0);
    {
        long long acc = 0;
        int q = 1;
        for (size_t t = 0; t < items.size() && q < NXCD; t++) {
            acc += cost[t];
            while (q < NXCD && acc * NXCD >= total * q) qbeg[(size_t)q++] = (int)t + 1;
        }
        while (q < NXCD) qbeg[(size_t)q++] = (int)items.size();
        qbeg[NXCD] = (int)items.size();
    }
    // (re)allocate and upload
    const size_t mb = sizeof(BMember<R>) * (size_t)nmem;
    if (b->mem_bytes < mb || b->items_cap < items.size() || b->nmem_cap < (size_t)nmem || b->fmap_cap < fmap.size() ||
        !b->d_qbeg) {
        free_tables(b);
        HIPCHK(hipMalloc(&b->d_mem, mb));
        HIPCHK(hipMalloc((void**)&b->d_items, sizeof(WorkItem) * std::max<size_t>(items.size(), 1)));
        HIPCHK(hipMalloc((void**)&b->d_qbeg, sizeof(int) * (NXCD + 1)));
        // two steps' queue counters, then the probe launch's done count
        HIPCHK(hipMalloc((void**)&b->d_ctr, sizeof(int) * (2 * NXCD + 1)));
        HIPCHK(hipMalloc((void**)&b->d_bstep, sizeof(long long)));
        HIPCHK(hipMalloc((void**)&b->d_fmap, sizeof(FaceBlock) * std::max<size_t>(fmap.size(), 1)));
        HIPCHK(hipMalloc((void**)&b->d_steps, sizeof(long long*) * (size_t)nmem));
        HIPCHK(hipMalloc((void**)&b->d_budget, sizeof(long long) * (size_t)nmem));
        b->mem_bytes = mb;
        b->items_cap = std::max<size_t>(items.size(), 1);
        b->nmem_cap = (size_t)nmem;
        b->fmap_cap = std::max<size_t>(fmap.size(), 1);
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(b->d_mem, mem.data(), mb, hipMemcpyHostToDevice));
    if (!items.empty())
        HIPCHK(hipMemcpy(b->d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_qbeg, qbeg.data(), sizeof(int) * (NXCD + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->d_ctr, 0, sizeof(int) * (2 * NXCD + 1)));
    HIPCHK(hipMemset(b->d_bstep, 0, sizeof(long long)));
    if (!fmap.empty())
        HIPCHK(hipMemcpy(b->d_fmap, fmap.data(), sizeof(FaceBlock) * fmap.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_steps, steps.data(), sizeof(long long*) * (size_t)nmem, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_budget, budget.data(), sizeof(long long) * (size_t)nmem, hipMemcpyHostToDevice));
    for (int i = 0; i < nmem; i++) {
        pf_ctx* c = b->members[(size_t)i];
        HIPCHK(hipMemcpy(c->d_step, &c->step, sizeof(long long), hipMemcpyHostToDevice));
    }
    b->face_blocks = fblk[(size_t)nmem];
    int rc = b->sponge == 1 ? sweep_grid_s<R, MODE, true>(b, tj, tz) : sweep_grid_s<R, MODE, false>(b, tj, tz);
    if (rc) return rc;
    if ((rc = free_graph(b))) return rc;
    b->dirty = false;
    return PF_OK;
}

int build(pf_batch* b) {
    switch (b->precision) {
        case PF_FP64: return build_t<double, MODE_FP64>(b);
        case PF_FP32: return build_t<float, MODE_FP32>(b);
        default: return build_t<float, MODE_FP32C>(b);
    }
}

template <class R, int MODE, bool SPONGE, int TJ, int KZ>
void launch_sweep(pf_batch* b) {
    hipLaunchKernelGGL((k_batch_sweep<R, MODE, SPONGE, TJ, KZ>), dim3((unsigned)b->grid), dim3(64, TJ * KZ / 64), 0,
                       b->stream, (const BMember<R>*)b->d_mem, (const WorkItem*)b->d_items, (const int*)b->d_qbeg,
                       b->d_ctr, (const long long*)b->d_bstep);
}

template <class R, int MODE, bool SPONGE>
void launch_sweep_s(pf_batch* b, int tj, int tz) {
    if (tz == 64 && tj == 16) launch_sweep<R, MODE, SPONGE, 16, 64>(b);
    else if (tz == 64 && tj == 8) launch_sweep<R, MODE, SPONGE, 8, 64>(b);
    else if (tz == 32 && tj == 32) launch_sweep<R, MODE, SPONGE, 32, 32>(b);
    else if (tz == 32 && tj == 16) launch_sweep<R, MODE, SPONGE, 16, 32>(b);
    else launch_sweep<R, MODE, SPONGE, 8, 32>(b);
}

// Issue one batched step (three launches) on the batch stream.
template <class R, int MODE>
int issue_step_t(pf_batch* b) {
    int tj, tz, tx;
    default_shape(b, tj, tz, tx);
    const int nmem = (int)b->members.size();
    if (b->sponge == 1) launch_sweep_s<R, MODE, true>(b, tj, tz);
    else launch_sweep_s<R, MODE, false>(b, tj, tz);
    if (b->face_blocks > 0)
        hipLaunchKernelGGL((k_batch_faces<R, MODE>), dim3((unsigned)b->face_blocks), dim3(256), 0, b->stream,
                           (const BMember<R>*)b->d_mem, (const FaceBlock*)b->d_fmap);
    hipLaunchKernelGGL((k_batch_probe<R>), dim3((unsigned)nmem), dim3(256), 0, b->stream, (const BMember<R>*)b->d_mem,
                       (long long* const*)b->d_steps, (const long long*)b->d_budget, nmem, b->d_bstep,
                       b->d_ctr + 2 * NXCD);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

int issue_step(pf_batch* b) {
    switch (b->precision) {
        case PF_FP64: return issue_step_t<double, MODE_FP64>(b);
        case PF_FP32: return issue_step_t<float, MODE_FP32>(b);
        default: return issue_step_t<float, MODE_FP32C>(b);
    }
}

// Capture k steps in a graph (replacing any previous one).
int capture(pf_batch* b, int k) {
    int rc = free_graph(b);
    if (rc) return rc;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < k && !rc; s++) rc = issue_step(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(PF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_steps = k;
    return PF_OK;
}

// Steps each member takes of the next n batched steps, and its new state.
void account(pf_batch* b, long long n) {
    for (pf_ctx* c : b->members) {
        const long long left = c->wave_len - c->step;
        const long long take = n < left ? n : (left > 0 ? left : 0);
        c->step += take;
        if (take & 1) {
            const int t = c->cur;
            c->cur = c->old;
            c->old = t;
        }
    }
}

constexpr int GRAPH_STEPS = 16;

int step_n(pf_batch* b, long long n) {
    int rc;
    if (b->dirty && (rc = build(b))) return rc;
    long long done = 0;
    if (b->use_graph && n >= GRAPH_STEPS) {
        if (b->graph_steps != GRAPH_STEPS && (rc = capture(b, GRAPH_STEPS))) return rc;
        for (; done + GRAPH_STEPS <= n; done += GRAPH_STEPS) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    }
    for (; done < n; done++)
        if ((rc = issue_step(b))) return rc;
    account(b, n);
    return PF_OK;
}

}  // namespace

// ------------------------------------------------------------ C ABI ----
extern "C" int pf_batch_create(int32_t precision, pf_batch** out) {
    if (!out) return fail(PF_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (precision < PF_FP64 || precision > PF_FP32C) return fail(PF_ERR_ARG, "bad precision %d", precision);
    pf_batch* b = new (std::nothrow) pf_batch();
    if (!b) return fail(PF_ERR_NOMEM, "out of host memory");
    b->precision = precision;
    int rc = PF_OK;
    if (hipGetDevice(&b->dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->dev) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(PF_ERR_HIP, "batch stream creation failed");
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return PF_OK;
}

extern "C" int pf_batch_add(pf_batch* b, pf_ctx* c) {
    if (!b || !c) return fail(PF_ERR_ARG, "NULL argument");
    if (c->batch) return fail(PF_ERR_STATE, "context is already in a batch");
    if (c->precision != b->precision) return fail(PF_ERR_ARG, "context precision %d != batch precision %d", c->precision, b->precision);
    if (c->nmat > PF_TILED_MAX_MATERIALS)
        return fail(PF_ERR_ARG, "batched sweeps hold at most %d materials (got %d)", PF_TILED_MAX_MATERIALS, c->nmat);
    if (level_bytes(c->cells, c->precision) >= (1ULL << 32))
        return fail(PF_ERR_ARG, "a batch member's time level must be < 4 GiB");
    const int sp = c->sponge ? 1 : 0;
    if (!b->members.empty() && b->sponge != sp)
        return fail(PF_ERR_ARG, "all members of a batch either have an absorbing layer or none");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->sponge = sp;
    b->members.push_back(c);
    c->batch = b;
    c->batch_stream = b->stream;
    b->dirty = true;
    return PF_OK;
}

extern "C" int pf_batch_remove(pf_batch* b, pf_ctx* c) {
    if (!b || !c) return fail(PF_ERR_ARG, "NULL argument");
    auto it = std::find(b->members.begin(), b->members.end(), c);
    if (it == b->members.end()) return fail(PF_ERR_STATE, "context is not a member of this batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->members.erase(it);
    c->batch = nullptr;
    c->batch_stream = nullptr;
    b->dirty = true;
    return PF_OK;
}

extern "C" int pf_batch_step(pf_batch* b, int64_t nsteps) {
    if (!b) return fail(PF_ERR_ARG, "batch is NULL");
    if (nsteps < 0) return fail(PF_ERR_ARG, "nsteps < 0");
    if (b->members.empty() || nsteps == 0) return PF_OK;
    return step_n(b, nsteps);
}

extern "C" int pf_batch_sync(pf_batch* b) {
    if (!b) return fail(PF_ERR_ARG, "batch is NULL");
    HIPCHK(hipStreamSynchronize(b->stream));
    return PF_OK;
}

extern "C" int pf_batch_set_tiles(pf_batch* b, int32_t tile_j, int32_t tile_z, int32_t tile_x, int32_t use_graph) {
    if (!b) return fail(PF_ERR_ARG, "batch is NULL");
    int tj = tile_j, tz = tile_z;
    if (tz == 0) tz = b->precision == PF_FP64 ? 32 : 64;
    if (tj == 0) tj = b->precision == PF_FP64 ? 8 : (tz == 64 ? 16 : 32);
    if (!shape_built(tj, tz)) return fail(PF_ERR_ARG, "batch tile shape %d rows x %d lanes is not built", tj, tz);
    if (tile_x < 0 || tile_x > XT - 3) return fail(PF_ERR_ARG, "tile_x must be 0..%d", XT - 3);
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tile_j = tile_j;
    b->tile_z = tile_z;
    b->tile_x = tile_x;
    b->use_graph = use_graph != 0;
    b->dirty = true;
    return PF_OK;
}

extern "C" int pf_batch_autotune(pf_batch* b, int32_t trial_steps, int32_t* tile_j, int32_t* tile_z, int32_t* tile_x,
                                 double* ms_per_step) {
    if (!b) return fail(PF_ERR_ARG, "batch is NULL");
    if (trial_steps < 1) return fail(PF_ERR_ARG, "trial_steps < 1");
    if (b->members.empty()) return fail(PF_ERR_STATE, "batch has no members");
    struct Cand { int tj, tz, tx; };
    std::vector<Cand> cands;
    const int xs[] = {16, 32, 64, 128};
    if (b->precision == PF_FP64) {
        for (int tx : xs) { cands.push_back({8, 32, tx}); cands.push_back({16, 32, tx}); cands.push_back({8, 64, tx}); }
    } else {
        for (int tx : xs) {
            cands.push_back({16, 64, tx}); cands.push_back({8, 64, tx});
            cands.push_back({32, 32, tx}); cands.push_back({16, 32, tx}); cands.push_back({8, 32, tx});
        }
    }
    // every candidate runs 1 untimed + trial_steps timed steps of real work
    long long room = 0;
    for (pf_ctx* c : b->members) room = std::max(room, c->wave_len - c->step);
    if (room < (long long)cands.size() * (trial_steps + 1))
        return fail(PF_ERR_RANGE, "members have %lld steps left; autotuning needs %lld", room,
                    (long long)cands.size() * (trial_steps + 1));
    const int keep_j = b->tile_j, keep_z = b->tile_z, keep_x = b->tile_x;
    const bool keep_g = b->use_graph;
    double best = 1e300;
    Cand bc = cands[0];
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, PF_TIMING_EVENT));
    HIPCHK(hipEventCreateWithFlags(&e1, PF_TIMING_EVENT));
    int rc = PF_OK;
    for (const Cand& k : cands) {
        b->tile_j = k.tj; b->tile_z = k.tz; b->tile_x = k.tx;
        b->use_graph = false;
        b->dirty = true;
        if ((rc = step_n(b, 1))) break;  // builds the tables, warms the kernel
        if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = fail(PF_ERR_HIP, "event record"); break; }
        if ((rc = step_n(b, trial_steps))) break;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(PF_ERR_HIP, "event sync");
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / trial_steps;
        if (per < best) { best = per; bc = k; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        b->tile_j = keep_j; b->tile_z = keep_z; b->tile_x = keep_x;
        b->use_graph = keep_g;
        b->dirty = true;
        return rc;
    }
    b->tile_j = bc.tj; b->tile_z = bc.tz; b->tile_x = bc.tx;
    b->use_graph = keep_g;
    b->dirty = true;
    if (tile_j) *tile_j = bc.tj;
    if (tile_z) *tile_z = bc.tz;
    if (tile_x) *tile_x = bc.tx;
    if (ms_per_step) *ms_per_step = best;
    return PF_OK;
}

extern "C" int pf_batch_destroy(pf_batch* b) {
    if (!b) return PF_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (pf_ctx* c : b->members) {
        c->batch = nullptr;
        c->batch_stream = nullptr;
    }
    (void)free_graph(b);
    free_tables(b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PF_OK;
}
