// Diagnostics with no counterpart in the reference: a streaming kernel of
// known traffic, used to measure the device's practical HBM bandwidth for a
// given read/write mix (reported beside the vendor peak) and to calibrate
// hardware traffic counters (FETCH_SIZE / WRITE_SIZE per known byte).
#include <hip/hip_runtime.h>

#include <vector>

#include "pf_runtime.hpp"

using namespace pf;

namespace {

constexpr int MAXS = 8;  // streams of each kind

typedef float f4 __attribute__((ext_vector_type(4)));

struct Streams {
    const f4* in[MAXS];
    f4* out[MAXS];
};

// One pass: block b owns the U * blockDim consecutive float4 elements from
// b * U * blockDim on; thread t handles elements t, t + blockDim, ...  All
// of a thread's loads are issued before the first use, so U * R loads are
// in flight per thread.  NT: non-temporal (streaming) loads and stores.
// Every index is checked against n4.
template <int R, int W, int U, bool NT>
__global__ void k_stream(Streams s, long long n4) {
    const long long base = (long long)blockIdx.x * blockDim.x * U + threadIdx.x;
    f4 v[U][R];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long i = base + (long long)u * blockDim.x;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (i < n4) v[u][r] = NT ? __builtin_nontemporal_load(&s.in[r][i]) : s.in[r][i];
            else v[u][r] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long i = base + (long long)u * blockDim.x;
        f4 a = v[u][0];
#pragma unroll
        for (int r = 1; r < R; r++) a += v[u][r];
        if (i < n4) {
#pragma unroll
            for (int w = 0; w < W; w++) {
                if (NT) __builtin_nontemporal_store(a, &s.out[w][i]);
                else s.out[w][i] = a;
            }
        }
    }
}

typedef void (*LaunchFn)(const Streams&, long long, hipStream_t);

template <int R, int W, int U, bool NT, int BS>
void launch(const Streams& s, long long n4, hipStream_t st) {
    const long long blocks = (n4 + (long long)BS * U - 1) / ((long long)BS * U);
    hipLaunchKernelGGL((k_stream<R, W, U, NT>), dim3((unsigned)blocks), dim3(BS), 0, st, s, n4);
}

// The launch shapes tried for a mix (the best one's rate is reported):
// one or two float4 per stream and thread, 256 or 512 threads, plain or
// non-temporal accesses.
template <int R, int W>
void shapes(std::vector<LaunchFn>& v) {
    v.push_back(launch<R, W, 1, false, 256>);
    v.push_back(launch<R, W, 1, false, 512>);
    v.push_back(launch<R, W, 2, false, 256>);
    v.push_back(launch<R, W, 2, true, 512>);
    v.push_back(launch<R, W, 1, true, 256>);
}

bool pick(int r, int w, std::vector<LaunchFn>& v) {
    if (r == 1 && w == 1) shapes<1, 1>(v);
    else if (r == 6 && w == 3) shapes<6, 3>(v);
    else if (r == 7 && w == 3) shapes<7, 3>(v);
    else if (r == 2 && w == 1) shapes<2, 1>(v);
    else return false;
    return true;
}

}  // namespace

extern "C" int pf_stream_probe(int32_t reads, int32_t writes, int64_t n, int32_t reps, double* gbs, double* ms) {
    if (!gbs || !ms) return fail(PF_ERR_ARG, "null output");
    std::vector<LaunchFn> fns;
    if (!pick(reads, writes, fns))
        return fail(PF_ERR_ARG, "stream mix %d reads / %d writes is not built (1/1, 2/1, 6/3, 7/3)", reads, writes);
    if (n <= 0 || n % 4 || reps < 1) return fail(PF_ERR_ARG, "n must be a positive multiple of 4, reps >= 1");
    if (n / 4 > 0x7fffffffLL * 256) return fail(PF_ERR_ARG, "n too large");
    Streams s = {};
    std::vector<void*> bufs;
    int rc = PF_OK;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    auto cleanup = [&]() {
        for (void* p : bufs) (void)hipFree(p);
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (st) (void)hipStreamDestroy(st);
    };
    for (int i = 0; i < reads + writes; i++) {
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)n * sizeof(float)) != hipSuccess) {
            cleanup();
            return fail(PF_ERR_NOMEM, "stream probe: cannot allocate %lld floats", (long long)n);
        }
        bufs.push_back(p);
        if (hipMemset(p, 0, (size_t)n * sizeof(float)) != hipSuccess) { cleanup(); return fail(PF_ERR_HIP, "memset"); }
        if (i < reads) s.in[i] = (const f4*)p; else s.out[i - reads] = (f4*)p;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&a, PF_TIMING_EVENT) != hipSuccess ||
        hipEventCreateWithFlags(&b, PF_TIMING_EVENT) != hipSuccess) {
        cleanup();
        return fail(PF_ERR_HIP, "stream probe: device set-up failed");
    }
    const long long n4 = n / 4;
    double best = 0.0;
    for (LaunchFn fn : fns) {
        fn(s, n4, st);  // warm-up
        fn(s, n4, st);
        (void)hipEventRecord(a, st);
        for (int r = 0; r < reps; r++) fn(s, n4, st);
        (void)hipEventRecord(b, st);
        float t = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess ||
            hipGetLastError() != hipSuccess) {
            rc = fail(PF_ERR_HIP, "stream probe: launch failed");
            break;
        }
        const double ms1 = (double)t / reps;
        const double g = (double)n * sizeof(float) * (reads + writes) / (ms1 * 1e-3) / 1e9;
        if (g > best) {
            best = g;
            *ms = ms1;
            *gbs = g;
        }
    }
    cleanup();
    return rc;
}
