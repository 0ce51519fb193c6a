// Spectral analysis of recorded traces: the transforms of the reference's
// analysis.spectrum (phonomena/simulation/analysis.py:44-96), which takes
// |fft(u * hann, norm="ortho")| of a time trace, or |fft2| of an (x, t)
// slice, and keeps the first N//2 frequencies.
//
// pf_dft: a batch of 1-D discrete Fourier transforms along the contiguous
// axis, in FP64, for the first nf frequencies:
//     out[b, k] = scale * sum_t (in[b, t] * w[t]) e^{-2 pi i k t / n}
// One thread per (b, k).  The trace row is the same for all threads of a
// block (a broadcast read); the twiddle e^{-2 pi i m / n} comes from a table
// of the n roots computed on the host in FP64, indexed by k*t mod n (kept
// incrementally, so no product overflows and no angle loses precision).
// The sum is direct, O(n) per output: traces here are a few thousand
// frames, and one launch covers every trace of a probe grid.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/phonofdtd.h"
#include "pf_runtime.hpp"

using namespace pf;

namespace {

__global__ __launch_bounds__(256) void k_dft(const double* __restrict__ re, const double* __restrict__ im,
                                             const double* __restrict__ win, long long n, long long nf,
                                             const double2* __restrict__ tw, double scale, double* __restrict__ ore,
                                             double* __restrict__ oim) {
    const long long b = blockIdx.y;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nf) return;
    const double* xr = re + b * n;
    const double* xi = im ? im + b * n : nullptr;
    double ar = 0.0, ai = 0.0;
    long long m = 0;  // k * t mod n
    for (long long t = 0; t < n; t++) {
        double x = xr[t], y = xi ? xi[t] : 0.0;
        if (win) {
            x = x * win[t];
            y = y * win[t];
        }
        const double2 c = tw[m];  // (cos, sin) of 2 pi m / n
        ar = ar + (x * c.x + y * c.y);
        ai = ai + (y * c.x - x * c.y);
        m += k;
        if (m >= n) m -= n;
    }
    ore[b * nf + k] = scale * ar;
    oim[b * nf + k] = scale * ai;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int pf_dft(const double* in_re, const double* in_im, int64_t batch, int64_t n, int64_t nf,
                      const double* window, double scale, double* out_re, double* out_im) {
    if (!in_re || !out_re || !out_im) return fail(PF_ERR_ARG, "NULL argument");
    if (batch < 1 || n < 1 || nf < 1 || nf > n) return fail(PF_ERR_ARG, "need batch, n >= 1 and 1 <= nf <= n");
    if (batch > 65535) return fail(PF_ERR_ARG, "batch must be <= 65535 (got %lld)", (long long)batch);
    std::vector<double2> tw((size_t)n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t m = 0; m < n; m++) {
        const double a = two_pi * (double)m / (double)n;
        tw[(size_t)m] = make_double2(std::cos(a), std::sin(a));
    }
    const size_t in_bytes = (size_t)batch * (size_t)n * sizeof(double);
    const size_t out_bytes = (size_t)batch * (size_t)nf * sizeof(double);
    DevBuf dre, dim, dwin, dtw, dore, doim;
    HIPCHK(hipMalloc(&dre.p, in_bytes));
    HIPCHK(hipMemcpy(dre.p, in_re, in_bytes, hipMemcpyHostToDevice));
    if (in_im) {
        HIPCHK(hipMalloc(&dim.p, in_bytes));
        HIPCHK(hipMemcpy(dim.p, in_im, in_bytes, hipMemcpyHostToDevice));
    }
    if (window) {
        HIPCHK(hipMalloc(&dwin.p, (size_t)n * sizeof(double)));
        HIPCHK(hipMemcpy(dwin.p, window, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&dtw.p, (size_t)n * sizeof(double2)));
    HIPCHK(hipMemcpy(dtw.p, tw.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&dore.p, out_bytes));
    HIPCHK(hipMalloc(&doim.p, out_bytes));
    const dim3 grid((unsigned)((nf + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(k_dft, grid, dim3(256), 0, nullptr, (const double*)dre.p, (const double*)dim.p,
                       (const double*)dwin.p, (long long)n, (long long)nf, (const double2*)dtw.p, scale,
                       (double*)dore.p, (double*)doim.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out_re, dore.p, out_bytes, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out_im, doim.p, out_bytes, hipMemcpyDeviceToHost));
    return PF_OK;
}
