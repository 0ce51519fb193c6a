// Spectral analysis of recorded traces: the transforms of the reference's
// analysis.spectrum (phonomena/simulation/analysis.py:44-96), which takes
// |fft(u * hann, norm="ortho")| of a time trace, or |fft2| of an (x, t)
// slice, and keeps the first N//2 frequencies.
//
// pf_dft: a batch of 1-D discrete Fourier transforms along the contiguous
// axis, in FP64, for the first nf frequencies:
//     out[b, k] = scale * sum_t (in[b, t] * w[t]) e^{-2 pi i k t / n}
// in O(n log n) per trace, for any n:
//
//  * n a power of two up to 8192: a radix-2 FFT of length M = n.
//  * any other n (up to 2^23): Bluestein's chirp-z form.  With c_m = e^{-pi i m^2 / n},
//    k t = (k^2 + t^2 - (k - t)^2) / 2 turns the transform into
//        X_k = c_k sum_t (a_t c_t) conj(c_{k-t}),
//    a convolution, evaluated as a circular one of power-of-two length
//    M >= 2n - 1 through two FFTs and the (precomputed) transform of the
//    wrapped chirp conj(c).  m^2 is reduced mod 2n in integers, so no
//    angle loses precision however long the trace.
//
// Every FFT is radix-2 in LDS: decimation in time on bit-reversed input
// (natural-order output) for the forward transforms, decimation in
// frequency (bit-reversed output) with conjugate twiddles for the inverse,
// so no pass reorders data.  M <= 8192 (128 KB): one block per trace does
// the whole transform in LDS.  Longer: M = M1 * M2 with t = M2 t1 + t2,
// k = k1 + M1 k2 (the four-step split), in three launches over a scratch of
// M values per trace:
//   1. length-M1 FFTs over t1 (several t2 columns per block, read as
//      contiguous runs), times e^{-2 pi i t2 k1 / M}, stored as [k1][t2];
//   2. per row k1: length-M2 FFT over t2 (now X at k1 + M1 k2, in row
//      order), times the chirp transform (kept in the same [k1][k2] order),
//      the inverse length-M2 FFT, times e^{+2 pi i k1 t2 / M};
//   3. inverse length-M1 FFTs over k1, times c_k * scale / M, the first nf
//      frequencies stored.
// The twiddles are host tables in FP64 (cos/sin of exact fractions of a
// turn); e^{-2 pi i m / M} for the four-step factor is the product of two
// table entries (m = a M2 + b).  Every product is plain FP64 arithmetic;
// the results differ from numpy's FFT by rounding only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/phonofdtd.h"
#include "pf_runtime.hpp"

using namespace pf;

namespace {

constexpr int FFT_LG_SMALL = 13;  // one block, whole transform in LDS: M <= 8192 (128 KB)
constexpr int FFT_LG_TILE = 12;   // four-step blocks hold 4096 values (64 KB)
constexpr int FFT_LG_MAX = 24;    // M <= 2^24: n <= 2^23 for Bluestein
constexpr int FFT_TILE_THREADS = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ unsigned brev(unsigned i, int bits) { return bits ? __brev(i) >> (32 - bits) : 0u; }

// G transforms of length L = 2^lg, each contiguous in s, in place.  tw holds
// e^{-2 pi i j / Lt} for j < Lt / 2, Lt = 2^tlg >= L; the butterfly of span
// h takes e^{-2 pi i j / 2h} = tw[j * Lt / 2h].  inv conjugates the twiddles.
// Both begin and end with a barrier.
__device__ void fft_dit(double2* s, int G, int lg, const double2* __restrict__ tw, int tlg, bool inv) {
    const int half = (1 << lg) >> 1, nb = G * half;
    for (int e = 0; e < lg; e++) {
        const int h = 1 << e;
        __syncthreads();
        for (int q = threadIdx.x; q < nb; q += blockDim.x) {
            const int g = q >> (lg - 1), r = q & (half - 1), j = r & (h - 1);
            double2* p = s + (g << lg) + ((r - j) << 1) + j;
            double2 w = tw[j << (tlg - e - 1)];
            if (inv) w.y = -w.y;
            const double2 a = p[0], b = cmul(p[h], w);
            p[0] = make_double2(a.x + b.x, a.y + b.y);
            p[h] = make_double2(a.x - b.x, a.y - b.y);
        }
    }
    __syncthreads();
}
// Natural-order input, bit-reversed output.
__device__ void fft_dif(double2* s, int G, int lg, const double2* __restrict__ tw, int tlg, bool inv) {
    const int half = (1 << lg) >> 1, nb = G * half;
    for (int e = lg - 1; e >= 0; e--) {
        const int h = 1 << e;
        __syncthreads();
        for (int q = threadIdx.x; q < nb; q += blockDim.x) {
            const int g = q >> (lg - 1), r = q & (half - 1), j = r & (h - 1);
            double2* p = s + (g << lg) + ((r - j) << 1) + j;
            double2 w = tw[j << (tlg - e - 1)];
            if (inv) w.y = -w.y;
            const double2 a = p[0], b = p[h];
            p[0] = make_double2(a.x + b.x, a.y + b.y);
            p[h] = cmul(make_double2(a.x - b.x, a.y - b.y), w);
        }
    }
    __syncthreads();
}

// Input value t of a trace: (re + i im) * win * chirp, zero past n.
__device__ __forceinline__ double2 load_in(const double* __restrict__ re, const double* __restrict__ im,
                                           const double* __restrict__ win, const double2* __restrict__ chirp,
                                           long long n, long long t) {
    double2 v = make_double2(0.0, 0.0);
    if (t < n) {
        v.x = re[t];
        v.y = im ? im[t] : 0.0;
        if (win) {
            v.x = v.x * win[t];
            v.y = v.y * win[t];
        }
        if (chirp) v = cmul(v, chirp[t]);
    }
    return v;
}

// e^{-2 pi i m / M}, m = a M2 + b, from the two tables.
__device__ __forceinline__ double2 twiddle_m(const double2* __restrict__ ta, const double2* __restrict__ tb,
                                             unsigned long long m, int lg2) {
    return cmul(ta[m >> lg2], tb[m & ((1ull << lg2) - 1)]);
}

// Whole transform of one trace (blockIdx.x) in LDS, M = 2^lg <= CAP.
//   bhat == nullptr, oc == nullptr: plain FFT (M = n), first nf values
//     times scale to ore/oim;
//   bhat != nullptr: Bluestein (chirp set), scale includes 1 / M;
//   oc != nullptr: forward FFT of the M input values, all stored to oc
//     (the chirp transform).
template <int CAP>
__global__ __launch_bounds__(1024) void k_fft_small(const double* __restrict__ re, const double* __restrict__ im,
                                                    const double* __restrict__ win,
                                                    const double2* __restrict__ chirp,
                                                    const double2* __restrict__ bhat, long long n, int lg,
                                                    const double2* __restrict__ tw, long long nf, double scale,
                                                    double* __restrict__ ore, double* __restrict__ oim,
                                                    double2* __restrict__ oc) {
    __shared__ double2 s[CAP];
    const int M = 1 << lg;
    const long long b = blockIdx.x;
    const double* rb = re + b * n;
    const double* ib = im ? im + b * n : nullptr;
    for (int t = threadIdx.x; t < M; t += blockDim.x) s[brev((unsigned)t, lg)] = load_in(rb, ib, win, chirp, n, t);
    fft_dit(s, 1, lg, tw, lg, false);
    if (oc) {
        for (int t = threadIdx.x; t < M; t += blockDim.x) oc[t] = s[t];
        return;
    }
    if (bhat) {
        for (int t = threadIdx.x; t < M; t += blockDim.x) s[t] = cmul(s[t], bhat[t]);
        fft_dif(s, 1, lg, tw, lg, true);
    }
    for (long long k = threadIdx.x; k < nf; k += blockDim.x) {
        double2 v = bhat ? cmul(s[brev((unsigned)k, lg)], chirp[k]) : s[k];
        ore[b * nf + k] = scale * v.x;
        oim[b * nf + k] = scale * v.y;
    }
}

// Four-step, launch 1: grid (M2 / cb, traces); block x takes the cb
// columns t2 = cb x ... cb x + cb - 1 of trace blockIdx.y.
__global__ __launch_bounds__(FFT_TILE_THREADS) void k_fft_cols(
    const double* __restrict__ re, const double* __restrict__ im, const double* __restrict__ win,
    const double2* __restrict__ chirp, long long n, int lg1, int lg2, int lgcb, const double2* __restrict__ tw,
    const double2* __restrict__ ta, const double2* __restrict__ tb, double2* __restrict__ T) {
    __shared__ double2 s[1 << FFT_LG_TILE];
    const int M1 = 1 << lg1, cb = 1 << lgcb;
    const long long b = blockIdx.y, M = 1ll << (lg1 + lg2);
    const int c0 = blockIdx.x << lgcb;
    const double* rb = re + b * n;
    const double* ib = im ? im + b * n : nullptr;
    for (int q = threadIdx.x; q < cb * M1; q += blockDim.x) {
        const int t1 = q >> lgcb, c = q & (cb - 1);
        s[(c << lg1) + brev((unsigned)t1, lg1)] = load_in(rb, ib, win, chirp, n, ((long long)t1 << lg2) + c0 + c);
    }
    fft_dit(s, cb, lg1, tw, lg2, false);
    double2* Tb = T + b * M;
    for (int q = threadIdx.x; q < cb * M1; q += blockDim.x) {
        const int k1 = q >> lgcb, c = q & (cb - 1), t2 = c0 + c;
        const unsigned long long m = ((unsigned long long)t2 * (unsigned)k1) & (unsigned long long)(M - 1);
        Tb[((long long)k1 << lg2) + t2] = cmul(s[(c << lg1) + k1], twiddle_m(ta, tb, m, lg2));
    }
}

// Launch 2: grid (M1 / rb, traces); rows k1 = rb x ... of trace blockIdx.y.
// bhatp == nullptr: forward rows only (the chirp transform's own set-up).
__global__ __launch_bounds__(FFT_TILE_THREADS) void k_fft_rows(double2* __restrict__ T,
                                                               const double2* __restrict__ bhatp, int lg1,
                                                               int lg2, int lgrb, const double2* __restrict__ tw,
                                                               const double2* __restrict__ ta,
                                                               const double2* __restrict__ tb) {
    __shared__ double2 s[1 << FFT_LG_TILE];
    const int M2 = 1 << lg2, cnt = M2 << lgrb;
    const long long b = blockIdx.y, M = 1ll << (lg1 + lg2);
    const int k10 = blockIdx.x << lgrb;
    double2* row = T + b * M + ((long long)k10 << lg2);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int r = q >> lg2, t2 = q & (M2 - 1);
        s[(r << lg2) + brev((unsigned)t2, lg2)] = row[q];
    }
    fft_dit(s, 1 << lgrb, lg2, tw, lg2, false);
    if (!bhatp) {
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) row[q] = s[q];
        return;
    }
    const double2* bh = bhatp + ((long long)k10 << lg2);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) s[q] = cmul(s[q], bh[q]);
    fft_dif(s, 1 << lgrb, lg2, tw, lg2, true);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int r = q >> lg2, t2 = q & (M2 - 1), k1 = k10 + r;
        const unsigned long long m = ((unsigned long long)t2 * (unsigned)k1) & (unsigned long long)(M - 1);
        row[q] = cmulc(s[(r << lg2) + brev((unsigned)t2, lg2)], twiddle_m(ta, tb, m, lg2));
    }
}

// Launch 3: grid (M2 / cb, traces); inverse columns, chirp, first nf out.
__global__ __launch_bounds__(FFT_TILE_THREADS) void k_fft_cols_out(const double2* __restrict__ T,
                                                                   const double2* __restrict__ chirp, long long nf,
                                                                   int lg1, int lg2, int lgcb,
                                                                   const double2* __restrict__ tw, double scale,
                                                                   double* __restrict__ ore,
                                                                   double* __restrict__ oim) {
    __shared__ double2 s[1 << FFT_LG_TILE];
    const int M1 = 1 << lg1, cb = 1 << lgcb;
    const long long b = blockIdx.y, M = 1ll << (lg1 + lg2);
    const int c0 = blockIdx.x << lgcb;
    const double2* Tb = T + b * M;
    for (int q = threadIdx.x; q < cb * M1; q += blockDim.x) {
        const int k1 = q >> lgcb, c = q & (cb - 1);
        s[(c << lg1) + k1] = Tb[((long long)k1 << lg2) + c0 + c];
    }
    fft_dif(s, cb, lg1, tw, lg2, true);
    for (int q = threadIdx.x; q < cb * M1; q += blockDim.x) {
        const int t1 = q >> lgcb, c = q & (cb - 1);
        const long long k = ((long long)t1 << lg2) + c0 + c;
        if (k < nf) {
            const double2 v = cmul(s[(c << lg1) + brev((unsigned)t1, lg1)], chirp[k]);
            ore[b * nf + k] = scale * v.x;
            oim[b * nf + k] = scale * v.y;
        }
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// e^{-2 pi i num / den} for 0 <= num < den (exact fractions of a turn).
double2 unit(long long num, long long den) {
    const double a = 6.283185307179586476925286766559 * ((double)num / (double)den);
    return make_double2(std::cos(a), -std::sin(a));
}

int upload(DevBuf& d, const void* host, size_t bytes) {
    HIPCHK(hipMalloc(&d.p, bytes));
    HIPCHK(hipMemcpy(d.p, host, bytes, hipMemcpyHostToDevice));
    return PF_OK;
}

int ceil_lg(long long v) {
    int l = 0;
    while ((1ll << l) < v) l++;
    return l;
}

int small_launch(int lg, long long blocks, const double* re, const double* im, const double* win,
                 const double2* chirp, const double2* bhat, long long n, const double2* tw, long long nf,
                 double scale, double* ore, double* oim, double2* oc) {
    const int M = 1 << lg;
    const unsigned threads = (unsigned)std::min(1024, std::max(64, M / 4));
    const dim3 grid((unsigned)blocks);
    if (lg <= 8)
        hipLaunchKernelGGL(k_fft_small<256>, grid, dim3(threads), 0, nullptr, re, im, win, chirp, bhat, n, lg, tw,
                           nf, scale, ore, oim, oc);
    else if (lg <= 11)
        hipLaunchKernelGGL(k_fft_small<2048>, grid, dim3(threads), 0, nullptr, re, im, win, chirp, bhat, n, lg,
                           tw, nf, scale, ore, oim, oc);
    else
        hipLaunchKernelGGL(k_fft_small<(1 << FFT_LG_SMALL)>, grid, dim3(threads), 0, nullptr, re, im, win, chirp,
                           bhat, n, lg, tw, nf, scale, ore, oim, oc);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

}  // namespace

extern "C" int pf_dft(const double* in_re, const double* in_im, int64_t batch, int64_t n, int64_t nf,
                      const double* window, double scale, double* out_re, double* out_im) {
    if (!in_re || !out_re || !out_im) return fail(PF_ERR_ARG, "NULL argument");
    if (batch < 1 || n < 1 || nf < 1 || nf > n) return fail(PF_ERR_ARG, "need batch, n >= 1 and 1 <= nf <= n");
    if (batch > 65535) return fail(PF_ERR_ARG, "batch must be <= 65535 (got %lld)", (long long)batch);
    // n <= 2^(FFT_LG_MAX-1): every length's chirp convolution fits 2^FFT_LG_MAX
    if (n > (1ll << (FFT_LG_MAX - 1)))
        return fail(PF_ERR_ARG, "n = %lld exceeds the longest transform (%lld)", (long long)n,
                    (1ll << (FFT_LG_MAX - 1)));
    // a power of two that fits one block: plain FFT of length n; otherwise
    // the chirp form over M >= 2n - 1 (the four-step path always takes it)
    const int lgn = ceil_lg(n);
    const bool blue = (1ll << lgn) != n || lgn > FFT_LG_SMALL;
    const int lg = blue ? ceil_lg(2 * n - 1) : lgn;
    const long long M = 1ll << lg;
    const bool small = lg <= FFT_LG_SMALL;
    const int lg1 = lg / 2, lg2 = lg - lg1;
    const int tlg = small ? lg : lg2;  // length of the longest FFT pass

    // host tables
    std::vector<double2> tw((size_t)std::max(1, (1 << tlg) / 2));
    for (long long j = 0; j < (1ll << tlg) / 2; j++) tw[(size_t)j] = unit(j, 1ll << tlg);
    std::vector<double2> chirp;
    std::vector<double> v_re, v_im;
    if (blue) {
        chirp.resize((size_t)n);
        for (long long m = 0; m < n; m++) {
            const long long r = (long long)(((unsigned long long)m * (unsigned long long)m) % (2ull * (unsigned long long)n));
            chirp[(size_t)m] = unit(r, 2 * n);  // e^{-pi i r / n}
        }
        // wrapped conj(c): v[m] = v[M - m] = conj(c_m), zero elsewhere
        v_re.assign((size_t)M, 0.0);
        v_im.assign((size_t)M, 0.0);
        for (long long m = 0; m < n; m++) {
            v_re[(size_t)m] = chirp[(size_t)m].x;
            v_im[(size_t)m] = -chirp[(size_t)m].y;
            if (m) {
                v_re[(size_t)(M - m)] = chirp[(size_t)m].x;
                v_im[(size_t)(M - m)] = -chirp[(size_t)m].y;
            }
        }
    }
    std::vector<double2> ta, tb;
    if (!small) {
        ta.resize((size_t)1 << lg1);
        tb.resize((size_t)1 << lg2);
        for (long long a = 0; a < (1ll << lg1); a++) ta[(size_t)a] = unit(a << lg2, M);
        for (long long c = 0; c < (1ll << lg2); c++) tb[(size_t)c] = unit(c, M);
    }

    const size_t in_bytes = (size_t)batch * (size_t)n * sizeof(double);
    const size_t out_bytes = (size_t)batch * (size_t)nf * sizeof(double);
    DevBuf dre, dim, dwin, dtw, dch, dvr, dvi, dbh, dta, dtb, dT, dore, doim;
    int rc;
    if ((rc = upload(dre, in_re, in_bytes))) return rc;
    if (in_im && (rc = upload(dim, in_im, in_bytes))) return rc;
    if (window && (rc = upload(dwin, window, (size_t)n * sizeof(double)))) return rc;
    if ((rc = upload(dtw, tw.data(), tw.size() * sizeof(double2)))) return rc;
    if (blue) {
        if ((rc = upload(dch, chirp.data(), chirp.size() * sizeof(double2)))) return rc;
        if ((rc = upload(dvr, v_re.data(), (size_t)M * sizeof(double)))) return rc;
        if ((rc = upload(dvi, v_im.data(), (size_t)M * sizeof(double)))) return rc;
        HIPCHK(hipMalloc(&dbh.p, (size_t)M * sizeof(double2)));
    }
    if (!small) {
        if ((rc = upload(dta, ta.data(), ta.size() * sizeof(double2)))) return rc;
        if ((rc = upload(dtb, tb.data(), tb.size() * sizeof(double2)))) return rc;
    }
    HIPCHK(hipMalloc(&dore.p, out_bytes));
    HIPCHK(hipMalloc(&doim.p, out_bytes));
    const double* re = (const double*)dre.p;
    const double* im = (const double*)dim.p;
    const double* win = (const double*)dwin.p;
    const double2* dtwp = (const double2*)dtw.p;
    const double2* ch = (const double2*)dch.p;
    double2* bh = (double2*)dbh.p;
    double* ore = (double*)dore.p;
    double* oim = (double*)doim.p;

    if (small) {
        if (blue) {
            if ((rc = small_launch(lg, 1, (const double*)dvr.p, (const double*)dvi.p, nullptr, nullptr, nullptr, M,
                                   dtwp, 0, 1.0, nullptr, nullptr, bh)))
                return rc;
        }
        if ((rc = small_launch(lg, batch, re, im, win, blue ? ch : nullptr, blue ? bh : nullptr, n, dtwp, nf,
                               blue ? scale / (double)M : scale, ore, oim, nullptr)))
            return rc;
    } else {
        const int lgcb = FFT_LG_TILE - lg1, lgrb = FFT_LG_TILE - lg2;  // columns / rows per block
        const double2* tap = (const double2*)dta.p;
        const double2* tbp = (const double2*)dtb.p;
        // chirp transform, in the [k1][k2] order the rows launch reads
        hipLaunchKernelGGL(k_fft_cols, dim3(1u << (lg2 - lgcb), 1), dim3(FFT_TILE_THREADS), 0, nullptr,
                           (const double*)dvr.p, (const double*)dvi.p, (const double*)nullptr,
                           (const double2*)nullptr, M, lg1, lg2, lgcb, dtwp, tap, tbp, bh);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_fft_rows, dim3(1u << (lg1 - lgrb), 1), dim3(FFT_TILE_THREADS), 0, nullptr, bh,
                           (const double2*)nullptr, lg1, lg2, lgrb, dtwp, tap, tbp);
        HIPCHK(hipGetLastError());
        // traces in chunks whose scratch stays within 1 GiB
        const long long chunk = std::max(1ll, std::min((long long)batch, (1ll << 30) / (M * 16)));
        HIPCHK(hipMalloc(&dT.p, (size_t)(chunk * M) * sizeof(double2)));
        double2* T = (double2*)dT.p;
        for (long long b0 = 0; b0 < batch; b0 += chunk) {
            const unsigned nb = (unsigned)std::min(chunk, (long long)batch - b0);
            hipLaunchKernelGGL(k_fft_cols, dim3(1u << (lg2 - lgcb), nb), dim3(FFT_TILE_THREADS), 0, nullptr,
                               re + b0 * n, im ? im + b0 * n : nullptr, win, ch, (long long)n, lg1, lg2, lgcb, dtwp,
                               tap, tbp, T);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_fft_rows, dim3(1u << (lg1 - lgrb), nb), dim3(FFT_TILE_THREADS), 0, nullptr, T,
                               (const double2*)bh, lg1, lg2, lgrb, dtwp, tap, tbp);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_fft_cols_out, dim3(1u << (lg2 - lgcb), nb), dim3(FFT_TILE_THREADS), 0, nullptr,
                               (const double2*)T, ch, (long long)nf, lg1, lg2, lgcb, dtwp, scale / (double)M,
                               ore + b0 * nf, oim + b0 * nf);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipMemcpy(out_re, dore.p, out_bytes, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out_im, doim.p, out_bytes, hipMemcpyDeviceToHost));
    return PF_OK;
}
