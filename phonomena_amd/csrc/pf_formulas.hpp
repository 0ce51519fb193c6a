// Per-value formulas of the FDTD step, shared by every sweep kernel.
//
// Each function maps already-loaded values to a result, so the naive and the
// tiled kernels run literally the same arithmetic.  Flavours (EXACT):
//
//  * 1 (FP64 parity mode): the expression order of the reference
//    (phonomena/simulation/base_solver.py:323-517), true division by the
//    spacings, no contraction (the library is built with -ffp-contract=off).
//    On the same inputs this reproduces the NumPy solver bit for bit.
//  * 2 (FP64 parity mode, every spacing a power of two): the same
//    expressions with each division by a spacing h taken as the product
//    with 1/h, which the spacing tables then hold.  1/h is exact for h = 2^k,
//    and x * 2^-k and x / 2^k are both the correctly rounded value of the
//    same real number (subnormal and overflowing results included), so this
//    is bit-identical to flavour 1, for the price of a multiply instead of a
//    correctly rounded division (about ten instructions in FP64).
//  * 0 (FP32 modes): divisions replaced by products with inverse spacings,
//    explicit fused multiply-adds.
//
// Material rows hold 16 values: C00 C01 C02 C10 C11 C12 C20 C21 C22 C33 C44
// C55 coef pad pad pad, with coef = dt^2 / rho (pre-folded on the host the way
// the reference computes `self.m.dt**2/self.m.P`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PF_DEV __device__ __forceinline__

namespace pf {

enum : int { M_C00 = 0, M_C01, M_C02, M_C10, M_C11, M_C12, M_C20, M_C21, M_C22,
             M_C33, M_C44, M_C55, M_COEF, MAT_STRIDE = 16 };

// MODE_FP64P: MODE_FP64 on a mesh whose spacings are all powers of two
// (flavour 2 above); the host picks it, with reciprocal spacing tables.
enum : int { MODE_FP64 = 0, MODE_FP32 = 1, MODE_FP32C = 2, MODE_FP64P = 3 };

// The formula flavour of a mode.
constexpr int exact_of(int mode) { return mode == MODE_FP64 ? 1 : (mode == MODE_FP64P ? 2 : 0); }

PF_DEV float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
PF_DEV double mad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a / h in flavour 1; a * h (h holding 1/h) in flavour 2.
template <int EXACT, class R>
PF_DEV R over(R a, R h) {
    if constexpr (EXACT == 2) return a * h;
    else return a / h;
}

// ---- normal stresses T1..T3 -------------------------------------------------
// Interior (base_solver.py:329-351): row a of C times the three normal
// differences, each divided by its staggered spacing.
template <int EXACT, class R>
PF_DEV void normal3(const R* c, R dux, R duy, R duz, R hx, R hy, R hz, R& t1, R& t2, R& t3) {
    if constexpr (EXACT) {
        t1 = over<EXACT>(c[M_C00] * dux, hx) + over<EXACT>(c[M_C01] * duy, hy) + over<EXACT>(c[M_C02] * duz, hz);
        t2 = over<EXACT>(c[M_C10] * dux, hx) + over<EXACT>(c[M_C11] * duy, hy) + over<EXACT>(c[M_C12] * duz, hz);
        t3 = over<EXACT>(c[M_C20] * dux, hx) + over<EXACT>(c[M_C21] * duy, hy) + over<EXACT>(c[M_C22] * duz, hz);
    } else {
        // hx, hy, hz are inverse spacings here
        R ex = dux * hx, ey = duy * hy, ez = duz * hz;
        t1 = mad(c[M_C00], ex, mad(c[M_C01], ey, c[M_C02] * ez));
        t2 = mad(c[M_C10], ex, mad(c[M_C11], ey, c[M_C12] * ez));
        t3 = mad(c[M_C20], ex, mad(c[M_C21], ey, c[M_C22] * ez));
    }
}

// Free surface k = 0 (base_solver.py:408-418): same rows with the first
// spacing of each axis; the one-sided uz difference is uz[..,0] - 0; T3 = 0.
template <int EXACT, class R>
PF_DEV void normal3_surface(const R* c, R dux, R duy, R uz0, R hx0, R hy0, R hz0, R& t1, R& t2, R& t3) {
    R duz = uz0 - R(0);
    if constexpr (EXACT) {
        t1 = over<EXACT>(c[M_C00] * dux, hx0) + over<EXACT>(c[M_C01] * duy, hy0) + over<EXACT>(c[M_C02] * duz, hz0);
        t2 = over<EXACT>(c[M_C10] * dux, hx0) + over<EXACT>(c[M_C11] * duy, hy0) + over<EXACT>(c[M_C12] * duz, hz0);
    } else {
        R ex = dux * hx0, ey = duy * hy0, ez = duz * hz0;
        t1 = mad(c[M_C00], ex, mad(c[M_C01], ey, c[M_C02] * ez));
        t2 = mad(c[M_C10], ex, mad(c[M_C11], ey, c[M_C12] * ez));
    }
    t3 = R(0);
}

// ---- shear stresses T4..T6 --------------------------------------------------
// c * (a / h1 + b / h2); the caller passes the spacings in the order the
// reference divides by them (interior: base_solver.py:353-372, surface with
// its swapped spacings: :420-433).
template <int EXACT, class R>
PF_DEV R shear(R c, R a, R h1, R b, R h2) {
    if constexpr (EXACT) {
        return c * (over<EXACT>(a, h1) + over<EXACT>(b, h2));
    } else {
        return c * mad(a, h1, b * h2);
    }
}

// ---- displacement divergence ------------------------------------------------
// (d1 / g1 + d2 / g2) + d3 / g3  (base_solver.py:444-446 and analogues)
template <int EXACT, class R>
PF_DEV R div3(R d1, R g1, R d2, R g2, R d3, R g3) {
    if constexpr (EXACT) {
        return over<EXACT>(d1, g1) + over<EXACT>(d2, g2) + over<EXACT>(d3, g3);
    } else {
        return mad(d1, g1, mad(d2, g2, d3 * g3));
    }
}

// Free-surface u_z bracket with the reference's precedence
// (base_solver.py:514-516): (d1/g1 + d2/g2) + T3[..,1] - T3[..,0]/fdz0.
template <int EXACT, class R>
PF_DEV R div3_uz_surface(R d1, R g1, R d2, R g2, R t3k1, R t3k0, R fz0) {
    if constexpr (EXACT) {
        return over<EXACT>(d1, g1) + over<EXACT>(d2, g2) + t3k1 - over<EXACT>(t3k0, fz0);
    } else {
        return mad(d1, g1, mad(d2, g2, t3k1)) - t3k0 * fz0;
    }
}

// ---- leapfrog ---------------------------------------------------------------
// u_new = 2 u - u_old + coef * div (base_solver.py:441-443).  With a damping
// term s > 0 (sponge extension): ((2u - (1-s) u_old) + coef*div) / (1+s);
// in FP32 the division is a product with the hardware reciprocal of 1+s
// (v_rcp_f32, within 1 ulp; one per node, shared by its three components:
// the same expression in every call, so the compiler forms it once), where
// a correctly rounded quotient costs about ten instructions and made the
// two-step pass 0.7 % slower.  At s = 0 both forms give the bits of the
// undamped update (the reciprocal of 1 is 1; 1 * x and x / 1 are exact).
template <int EXACT, bool SPONGE, class R>
PF_DEV R leap(R u, R uo, R coef, R dv, R s) {
    if constexpr (EXACT) {
        if constexpr (SPONGE) {
            // off the absorbing layer s = 0: (1 - 0) u_old = u_old and
            // x / (1 + 0) = x exactly, so the division is skipped there
            R r = 2 * u - (1 - s) * uo + coef * dv;
            if (s != R(0)) r = r / (1 + s);
            return r;
        }
        else return 2 * u - uo + coef * dv;
    } else {
        if constexpr (SPONGE) return mad(coef, dv, 2 * u - (1 - s) * uo) * __builtin_amdgcn_rcpf(1 + s);
        else return mad(coef, dv, 2 * u - uo);
    }
}

// ---- compensated (hi + lo) arithmetic for PF_FP32C --------------------------
// Storage of the lo part: after every normalising TwoSum |lo| <= ulp(hi)/2,
// so lo is kept as a signed 16-bit multiple of 2^-16 ulp(hi) (ulp(hi) =
// 2^(e-24) for hi = m 2^e, 0.5 <= |m| < 1): hi + lo carries 24 + 16
// significant bits in 6 bytes instead of 8.  Both functions are exact
// scalings (ldexp) around one round-to-nearest; hi = 0 has lo = 0.
typedef int16_t lo_t;

PF_DEV lo_t lo_encode(float hi, float lo) {
    const int e = __builtin_amdgcn_frexp_expf(hi);
    float q = __builtin_rintf(__builtin_amdgcn_ldexpf(lo, 40 - e));
    q = fminf(fmaxf(q, -32767.0f), 32767.0f);
    return (lo_t)(int)q;
}
PF_DEV float lo_decode(float hi, lo_t q) {
    return __builtin_amdgcn_ldexpf((float)q, __builtin_amdgcn_frexp_expf(hi) - 40);
}
// FP64 fields carry no lo part.
PF_DEV lo_t lo_encode(double, double) { return 0; }
PF_DEV double lo_decode(double, lo_t) { return 0.0; }

// Error-free transformations (Knuth TwoSum, FMA TwoProd); a value is the
// unevaluated sum hi + lo with |lo| <= ulp(hi)/2 after normalisation.
PF_DEV void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// U_new = 2 U - U_old + f (+ damping), with U = hi + lo.  The increment f is
// an ordinary FP32 value; what is compensated is the accumulation of the
// displacement itself, so the small per-step increments are never lost
// against the large running value.
template <bool SPONGE>
PF_DEV void leap_comp(float hi, float lo, float ohi, float olo, float f, float s, float& nhi, float& nlo) {
    if constexpr (!SPONGE) {
        float a, e;
        two_sum(2.0f * hi, -ohi, a, e);
        float low = ((2.0f * lo - olo) + e) + f;
        two_sum(a, low, nhi, nlo);
    } else {
        float q = 1.0f - s;
        float p = q * ohi;
        float pe = mad(q, ohi, -p);
        float plow = mad(q, olo, pe);
        float a, e;
        two_sum(2.0f * hi, -p, a, e);
        float low = ((2.0f * lo - plow) + e) + f;
        float nh, nl;
        two_sum(a, low, nh, nl);
        float r = __builtin_amdgcn_rcpf(1.0f + s);  // as in leap()
        float h2 = nh * r;
        float l2 = mad(nh, r, -h2) + nl * r;
        two_sum(h2, l2, nhi, nlo);
    }
}

// a + c * (b - d) for the Mur faces in compensated form; c = ch + cl.
PF_DEV void mur_comp(float ah, float al, float ch, float cl, float bh, float bl, float dh, float dl,
                     float& rh, float& rl) {
    float sh, se;
    two_sum(bh, -dh, sh, se);
    se += bl - dl;
    float ph = ch * sh;
    float pe = mad(ch, sh, -ph);
    pe = mad(ch, se, mad(cl, sh, pe));
    float th, te;
    two_sum(ah, ph, th, te);
    te += al + pe;
    two_sum(th, te, rh, rl);
}

}  // namespace pf
