// FP64 elementary functions for the delay samplers, sized for sampling:
// results within a few ulp of the correctly rounded value, in a fraction of the
// instructions of a general-purpose libm.  Each is written from the textbook
// method named on it; the test suite compares them with the host libm over
// their sampling domains.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcdp_log_table.hpp"

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

// The breakpoint table of log_tab (the lookups are per lane): kLogInvF and
// kLogHiLo from mcdp_log_table.hpp, written by gen_log_table.py.  Kernels
// that sample many activities per workgroup stage it in LDS; the others read
// it where it lies (global_log_table).
struct LogTable {
    const double *inv;    // [kLogTableSize]: 1/F_j rounded
    const double2 *hilo;  // [kLogTableSize]: log F_j = hi + lo, hi a multiple of 2^-32
};
__device__ __forceinline__ LogTable global_log_table() {
    return LogTable{kLogInvF, reinterpret_cast<const double2 *>(kLogHiLo)};
}
struct LogTableLds {
    double inv[kLogTableSize];
    double2 hilo[kLogTableSize];
    __device__ LogTable view() const { return LogTable{inv, hilo}; }
};
// Every thread of the workgroup calls it; the caller's next barrier makes the
// table visible to all waves.
__device__ __forceinline__ void stage_log_table(LogTableLds &t) {
    for (int i = threadIdx.x; i < kLogTableSize; i += blockDim.x) {
        t.inv[i] = kLogInvF[i];
        t.hilo[i] = double2{kLogHiLo[2 * i], kLogHiLo[2 * i + 1]};
    }
}

// Natural logarithm of a positive finite x, table-driven (P. T. P. Tang,
// "Table-driven implementation of the logarithm function in IEEE
// floating-point arithmetic", ACM TOMS 16(4), 1990):
// x = 2^k m with m in [1/2, 1); F = 1/2 + j/256 the breakpoint nearest m;
// log x = k ln2 + log F + log(1 + r) with r = (m - F)/F, |r| <= 2^-8.
//   * m - F is exact (Sterbenz); r = (m - F) * (1/F) carries two roundings,
//     i.e. ~2^-60 absolute;
//   * log(1 + r) = r + r^2 P(r), Taylor to r^7 (remainder < 2^-67);
//   * k ln2_hi + hi is exact (both multiples of 2^-32, sum below 2^10), and
//     the small terms k ln2_lo + lo + log(1 + r) are added to it last.
// At x = 1 (k = 1, F = 1/2, log F = -ln2_hi - ln2_lo) both parts cancel to 0;
// near 1 from below F = 1 and the result is log(1 + r) alone, so the
// relative error stays within an ulp there too.
__device__ __forceinline__ double log_tab(double x, const LogTable &t) {
    const double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
    const int k = __builtin_amdgcn_frexp_exp(x);
    const double jd = __builtin_rint(__builtin_fma(m, 256.0, -128.0));  // exact argument; j in [0, 128]
    const int j = (int)jd;
    const double F = __builtin_fma(jd, 1.0 / 256.0, 0.5);  // exact
    const double r = (m - F) * t.inv[j];
    const double2 lf = t.hilo[j];
    const double r2 = r * r;
    double p = 1.0 / 7.0;
    p = fma_k(p, r, -1.0 / 6.0);
    p = fma_k(p, r, 1.0 / 5.0);
    p = fma_k(p, r, -1.0 / 4.0);
    p = fma_k(p, r, 1.0 / 3.0);
    p = __builtin_fma(p, r, -0.5);
    const double log1p_r = __builtin_fma(r2, p, r);
    const double kd = (double)k;
    constexpr double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    const double hi = __builtin_fma(kd, ln2_hi, lf.x);
    const double lo = __builtin_fma(kd, ln2_lo, lf.y) + log1p_r;
    return hi + lo;
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
