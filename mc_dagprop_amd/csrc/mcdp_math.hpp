// FP64 elementary functions for the delay samplers, sized for sampling:
// results within a few ulp of the correctly rounded value, in a fraction of the
// instructions of a general-purpose libm.  Each is written from the textbook
// method named on it; the test suite compares them with the host libm over
// their sampling domains.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcdp {

// p * z + c for a polynomial coefficient c.  The compiler's accumulate form
// (v_fmac_f64) wants c in vector registers and rebuilds each 64-bit constant
// with two v_mov_b32 per use; here c is bound to a scalar register pair, so a
// Horner step is one v_fma_f64 (the pair is set up by scalar moves, which
// issue beside the vector pipeline).
__device__ __forceinline__ double fma_k(double p, double z, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(p), "v"(z), "s"(c));
    return r;
#else
    return __builtin_fma(p, z, c);
#endif
}

// Quotient a / b to ~1 ulp: hardware reciprocal estimate refined by two
// Newton-Raphson steps, then one residual correction of the quotient.
__device__ __forceinline__ double div_fast(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

// Natural logarithm of a positive normal x.
// x = 2^k m with m in [sqrt(1/2), sqrt(2)); with s = (m - 1) / (m + 1),
// log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ...), |s| <= 0.1716, so ten
// terms of the series reach 2^-53; log x = k ln 2 + log m with ln 2 split in a
// high part exact in k * ln2_hi and a low part.
__device__ __forceinline__ double log_fast(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    int k = __builtin_amdgcn_frexp_exp(x);
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m + m : m;
    k = lo ? k - 1 : k;
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = div_fast(f, m + 1.0);
    const double z = s * s;
    double p = 1.0 / 21.0;
    p = fma_k(p, z, 1.0 / 19.0);
    p = fma_k(p, z, 1.0 / 17.0);
    p = fma_k(p, z, 1.0 / 15.0);
    p = fma_k(p, z, 1.0 / 13.0);
    p = fma_k(p, z, 1.0 / 11.0);
    p = fma_k(p, z, 1.0 / 9.0);
    p = fma_k(p, z, 1.0 / 7.0);
    p = fma_k(p, z, 1.0 / 5.0);
    p = fma_k(p, z, 1.0 / 3.0);
    // log m = 2s + 2s z p; the correction term is < 1% of 2s.
    const double two_s = s + s;
    const double log_m = __builtin_fma(two_s * z, p, two_s);
    const double kd = (double)k;
    constexpr double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    return __builtin_fma(kd, ln2_hi, __builtin_fma(kd, ln2_lo, log_m));
}

// cos(2 pi u) for u in [0, 1], to within ~3e-16 absolute.
// Reduce to w in [0, 1/4] by the symmetries of cosine (period 1, even,
// cos(pi - a) = -cos a), then the Taylor polynomial of cos in a = 2 pi w,
// |a| <= pi/2, to degree 22 (remainder (pi/2)^24 / 24! < 1e-22).  Near
// a = pi/2 the result is small and its error is absolute, which is what the
// samplers need (they use cos as a factor of a normal deviate).
__device__ __forceinline__ double cos2pi_fast(double u) {
    double w = u - __builtin_rint(u);  // [-1/2, 1/2]
    w = __builtin_fabs(w);             // cos is even
    const bool neg = w > 0.25;         // cos(2pi w) = -cos(2pi (1/2 - w))
    w = neg ? 0.5 - w : w;             // [0, 1/4]
    constexpr double two_pi = 6.28318530717958647693;
    const double a = w * two_pi;
    const double a2 = a * a;
    // cos a = sum (-1)^n a^{2n} / (2n)!
    double c = -1.0 / 1.1240007277776077e21;            // -1/22!
    c = fma_k(c, a2, 1.0 / 2432902008176640000.0);  //  1/20!
    c = fma_k(c, a2, -1.0 / 6402373705728000.0);    // -1/18!
    c = fma_k(c, a2, 1.0 / 20922789888000.0);   //  1/16!
    c = fma_k(c, a2, -1.0 / 87178291200.0);     // -1/14!
    c = fma_k(c, a2, 1.0 / 479001600.0);        //  1/12!
    c = fma_k(c, a2, -1.0 / 3628800.0);         // -1/10!
    c = fma_k(c, a2, 1.0 / 40320.0);            //  1/8!
    c = fma_k(c, a2, -1.0 / 720.0);             // -1/6!
    c = fma_k(c, a2, 1.0 / 24.0);               //  1/4!
    c = __builtin_fma(c, a2, -0.5);                    // -1/2!
    c = __builtin_fma(c, a2, 1.0);
    return neg ? -c : c;
}

// e^x for x <= 0 down to about -745 (the only range the samplers use):
// x = k ln 2 + r with |r| <= ln2 / 2 (Cody-Waite split of ln 2), Taylor
// polynomial of e^r to degree 13 (remainder < 2^-56), times 2^k by ldexp.
__device__ __forceinline__ double exp_fast(double x) {
    constexpr double inv_ln2 = 1.44269504088896338700;
    constexpr double ln2_hi = 6.93147180369123816490e-01;
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    const double kd = __builtin_rint(x * inv_ln2);
    const double r = __builtin_fma(-kd, ln2_lo, __builtin_fma(-kd, ln2_hi, x));
    double p = 1.0 / 6227020800.0;                 // 1/13!
    p = fma_k(p, r, 1.0 / 479001600.0);    // 1/12!
    p = fma_k(p, r, 1.0 / 39916800.0);     // 1/11!
    p = fma_k(p, r, 1.0 / 3628800.0);      // 1/10!
    p = fma_k(p, r, 1.0 / 362880.0);       // 1/9!
    p = fma_k(p, r, 1.0 / 40320.0);        // 1/8!
    p = fma_k(p, r, 1.0 / 5040.0);         // 1/7!
    p = fma_k(p, r, 1.0 / 720.0);          // 1/6!
    p = fma_k(p, r, 1.0 / 120.0);          // 1/5!
    p = fma_k(p, r, 1.0 / 24.0);           // 1/4!
    p = fma_k(p, r, 1.0 / 6.0);            // 1/3!
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)kd);
}

}  // namespace mcdp
