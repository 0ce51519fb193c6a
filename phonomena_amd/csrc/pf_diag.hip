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

struct Streams {
    const float4* in[MAXS];
    float4* out[MAXS];
};

// Each thread reads one float4 of every input stream and writes their sum
// to every output stream (grid-stride; every access is in bounds).
template <int R, int W>
__global__ __launch_bounds__(256) void k_stream(Streams s, long long n4) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float4 v = s.in[r][i];
            a.x += v.x; a.y += v.y; a.z += v.z; a.w += v.w;
        }
#pragma unroll
        for (int w = 0; w < W; w++) s.out[w][i] = a;
    }
}

template <int R, int W>
void launch(const Streams& s, long long n4, int blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_stream<R, W>), dim3(blocks), dim3(256), 0, st, s, n4);
}

typedef void (*LaunchFn)(const Streams&, long long, int, hipStream_t);

LaunchFn pick(int r, int w) {
    if (r == 1 && w == 1) return launch<1, 1>;
    if (r == 6 && w == 3) return launch<6, 3>;
    if (r == 7 && w == 3) return launch<7, 3>;
    if (r == 2 && w == 1) return launch<2, 1>;
    if (r == 1 && w == 0) return launch<1, 0>;
    return nullptr;
}

}  // namespace

extern "C" int pf_stream_probe(int32_t reads, int32_t writes, int64_t n, int32_t reps, double* gbs, double* ms) {
    if (!gbs || !ms) return fail(PF_ERR_ARG, "null output");
    LaunchFn fn = pick(reads, writes);
    if (!fn) return fail(PF_ERR_ARG, "stream mix %d reads / %d writes is not built (1/1, 2/1, 6/3, 7/3, 1/0)", reads, writes);
    if (n <= 0 || n % 4 || reps < 1) return fail(PF_ERR_ARG, "n must be a positive multiple of 4, reps >= 1");
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
        if (i < reads) s.in[i] = (const float4*)p; else s.out[i - reads] = (float4*)p;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
        hipEventCreate(&b) != hipSuccess) {
        cleanup();
        return fail(PF_ERR_HIP, "stream probe: device set-up failed");
    }
    const int blocks = cus * 8;
    const long long n4 = n / 4;
    fn(s, n4, blocks, st);  // warm-up
    fn(s, n4, blocks, st);
    (void)hipEventRecord(a, st);
    for (int r = 0; r < reps; r++) fn(s, n4, blocks, st);
    (void)hipEventRecord(b, st);
    float t = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess || hipGetLastError() != hipSuccess)
        rc = fail(PF_ERR_HIP, "stream probe: launch failed");
    if (rc == PF_OK) {
        *ms = (double)t / reps;
        *gbs = (double)n * sizeof(float) * (reads + writes) / (*ms * 1e-3) / 1e9;
    }
    cleanup();
    return rc;
}
