// Counter-based delay sampling on gfx950.
//
// The reference draws every delay from one sequential Xoshiro256++ stream per
// run (_core.cpp:313, _custom_rng.hpp:551-600) through libstdc++'s
// exponential / gamma / discrete distributions (_core.cpp:78-141).  A GPU lane
// cannot share a sequential stream, so each draw here is a pure function of
// (seed, activity, attempt): Philox (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11) keyed by the activity index,
// with the run's seed and the attempt number as counter.  A duration therefore
// does not depend on which lane, tile, batch or launch computes it, and
// run_many(seeds)[i] equals run(seeds[i]) exactly.
//
// Distributions are matched in law, not stream: each method below samples
// exactly the distribution the reference samples.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcdp_math.hpp"
#include "mcdp_plan.hpp"

namespace mcdp {

// Read-only views in the constant address space (4): loads through them are
// known not to alias any store of the kernel, so wave-uniform ones are issued
// as scalar loads.  (hipcc's host pass parses kernel bodies too; there the
// view is a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T *;
#else
template <typename T>
using const_ptr = const T *;
#endif
template <typename T>
__device__ __forceinline__ const_ptr<T> as_const_space(const T *p) {
    return (const_ptr<T>)p;
}

// ---- Philox -----------------------------------------------------------------
// Multipliers and Weyl key increments as published with Random123.
constexpr uint32_t kPhiloxM4a = 0xD2511F53u, kPhiloxM4b = 0xCD9E8D57u;
constexpr uint32_t kPhiloxM2 = 0xD256D193u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

struct Words4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c in one instruction: gfx950's three-input bit operation, truth
// table 0x96 (set where an odd number of the inputs is set).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}
struct Words2 {
    uint32_t x, y;
};

// Philox-4x32-10.  A 32x32->64 product is one v_mad_u64_u32.
__device__ __forceinline__ Words4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint64_t p0 = (uint64_t)kPhiloxM4a * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM4b * c2;
        const uint32_t y0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t y2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = y0;
        c2 = y2;
    }
    return Words4{c0, c1, c2, c3};
}

// Philox-2x32-10.
__device__ __forceinline__ Words2 philox2x32_10(uint32_t c0, uint32_t c1, uint32_t k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) k += kPhiloxW0;
        const uint64_t p = (uint64_t)kPhiloxM2 * c0;
        const uint32_t y0 = xor3((uint32_t)(p >> 32), k, c1);
        c1 = (uint32_t)p;
        c0 = y0;
    }
    return Words2{c0, c1};
}

// ---- uniforms ---------------------------------------------------------------
// 53 random bits m = (hi >> 11) * 2^32 + lo; the fma is exact (m < 2^53).
__device__ __forceinline__ double bits53(uint32_t hi, uint32_t lo) {
    return __builtin_fma((double)(hi >> 11), 4294967296.0, (double)lo);
}
// (m + 1) 2^-53, in (0, 1]: safe argument for log.  Both fmas are exact:
// (lo + 1) 2^-53, then (hi >> 11) 2^-21 plus it, a multiple of 2^-53 <= 1.
__device__ __forceinline__ double u01_open0(uint32_t hi, uint32_t lo) {
    return __builtin_fma((double)(hi >> 11), 0x1p-21, __builtin_fma((double)lo, 0x1p-53, 0x1p-53));
}
// m 2^-53, in [0, 1).
__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) { return bits53(hi, lo) * 0x1p-53; }
// (w + 1/2) 2^-32, in (0, 1): 32-bit resolution, used for an angle and for
// the acceptance test of Marsaglia-Tsang.
__device__ __forceinline__ double u01_mid32(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

// The random words of one draw.  Philox is keyed by the activity index (the
// same for every lane of a wave, so the round keys live in scalar registers)
// and counts over the run's seed: activity a is stream a, the seed is the
// position in it, a tag word separates the draws one activity may need.  The
// seed is the C int seed's low 32 bits; the reference widens it by sign
// extension (_custom_rng.hpp:570), so the high word carries nothing more.
//   two words:  Philox-2x32-10(ctr = (seed, tag), key = a)
//   four words: Philox-4x32-10(ctr = (seed, tag, attempt, 0), key = (a, 0))
__device__ __forceinline__ Words2 draw_words2(uint32_t edge, uint32_t seed, uint32_t tag) {
    return philox2x32_10(seed, tag, edge);
}
__device__ __forceinline__ Words4 draw_words4(uint32_t edge, uint32_t seed, uint32_t attempt, uint32_t tag) {
    return philox4x32_10(seed, tag, attempt, 0u, edge, 0u);
}

// ---- per-kind draws of the multiplier X (or the absolute extra) -------------

// One Marsaglia-Tsang attempt for shape k' = d + 1/3 >= 1, scale 1
// (Marsaglia & Tsang, ACM TOMS 26(3), 2000).  The normal comes from one
// Box-Muller branch.  Returns false on rejection.
__device__ __forceinline__ bool mt_attempt(const Words4 &w, double d, double c, const LogTable &lt, double &x) {
    const double u1 = u01_open0(w.x, w.y);
    const double u2 = u01_mid32(w.z);
    const double u3 = u01_mid32(w.w);
    const double z = sqrt(-2.0 * log_tab(u1, lt)) * cos2pi_fast(u2);
    double v = 1.0 + c * z;
    if (!(v > 0.0)) return false;
    v = v * v * v;
    const double z2 = z * z;
    x = d * v;
    if (u3 < 1.0 - 0.0331 * (z2 * z2)) return true;
    return log_tab(u3, lt) < 0.5 * z2 + d * (1.0 - v + log_tab(v, lt));
}

// Gamma family with rejection: on the truncation bound max_scale (and, for
// Marsaglia-Tsang and the power proposal, on the method's own test), with the
// attempt number in the counter so every lane's draw is fixed by (seed, edge)
// alone.  Returns X (scale included).
__device__ __forceinline__ double gamma_rejection(uint32_t kind, uint32_t flags, const DistParams &D, uint32_t edge,
                                                  uint32_t seed, const LogTable &lt) {
    const double m = D.p[3];
    double X = 0.0;
    uint32_t j = 0;
    for (;;) {
        const Words4 w = draw_words4(edge, seed, j, kTagMain);
        bool ok = true;
        if (kind == kGammaErlang2) {  // sum of two unit exponentials
            X = -D.p[0] * log_tab(u01_open0(w.x, w.y) * u01_open0(w.z, w.w), lt);
        } else if (kind == kGammaHalf) {  // Gamma(1/2) = N(0,1)^2 / 2 = Exp(1) cos^2(2 pi U)
            const double cs = cos2pi_fast(u01_mid32(w.z));
            X = D.p[0] * (-log_tab(u01_open0(w.x, w.y), lt) * (cs * cs));
        } else if (kind == kGammaMT) {
            double y;
            ok = mt_attempt(w, D.p[1], D.p[2], lt, y);
            if (ok) {
                if (flags & kInfoBoost) {  // shape < 1: Gamma(k) = Gamma(k + 1) U^{1/k}
                    const Words2 b = draw_words2(edge, seed, (j << 8) | kTagBoost);
                    y *= exp_fast(log_tab(u01_open0(b.x, b.y), lt) * D.p[4]);
                }
                X = D.p[0] * y;
            }
        } else {  // kGammaPow: density x^{k-1} e^{-x} on [0, m'] by proposal
                  // m' U^{1/k}, accepted with probability e^{-x}
            const double y = D.p[1] * exp_fast(log_tab(u01_open0(w.x, w.y), lt) * D.p[2]);
            ok = u01_open0(w.z, w.w) <= exp_fast(-y);
            X = (y * D.p[0] > m) ? m : y * D.p[0];
        }
        if (ok && !(X > m)) break;
        if (++j >= kMaxAttempts) {
            X = m;
            break;
        }
    }
    return X;
}

// Draws the duration of one activity for N samples (a lane's samples): the
// distribution switch is taken once and each case runs the N independent
// draws as straight-line code, so their dependency chains interleave.
// `info`, `dmin` and `p0` come from the activity's wave-uniform record, `edge`
// is its index; kinds that need more parameters read them from `dists`; `lt`
// is the logarithm's breakpoint table in LDS.
// Each result is base + extra rounded exactly as the reference (_core.cpp:75,
// 88, 103, 125, 140, 328): extra = X * d (one rounding), then d + extra (one
// rounding); this file is compiled with -ffp-contract=off.  `alias` points at
// the alias records, in LDS or in the constant address space.
template <int N, typename AliasPtr>
__device__ __forceinline__ void draw_durations(uint32_t info, double dmin, double p0, uint32_t edge,
                                               const uint32_t (&seed)[N], AliasPtr alias,
                                               const_ptr<DistParams> dists, const LogTable &lt, double (&dur)[N]) {
    const uint32_t kind = info & kInfoKindMask;
    double X[N];
    switch (kind) {
        case kNone:
#pragma unroll
            for (int i = 0; i < N; ++i) dur[i] = dmin;
            return;
        case kEmpiricalAbs1:
#pragma unroll
            for (int i = 0; i < N; ++i) dur[i] = dmin + p0;
            return;
        case kEmpiricalAbs:
        case kEmpiricalRel: {
            // Vose alias table: one uniform picks the bin and the coin.
            const uint64_t packed = (uint64_t)__double_as_longlong(p0);  // (n, table_off)
            const int n = (int)(uint32_t)packed;
            const int table_off = (int)(uint32_t)(packed >> 32);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const Words2 w = draw_words2(edge, seed[i], kTagMain);
                const double xn = u01(w.x, w.y) * (double)n;
                int b = (int)xn;
                b = (b >= n) ? n - 1 : b;
                const double frac = xn - (double)b;
                const AliasRec rec = alias[table_off + b];
                X[i] = (frac < rec.prob) ? rec.v_self : rec.v_alias;
            }
            if (kind == kEmpiricalAbs) {
#pragma unroll
                for (int i = 0; i < N; ++i) dur[i] = dmin + X[i];
                return;
            }
            break;
        }
        case kConstant:
        case kEmpiricalRel1:
#pragma unroll
            for (int i = 0; i < N; ++i) X[i] = p0;
            break;
        case kExpInf:  // inversion: X = -mean log U, U in (0, 1]
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const Words2 w = draw_words2(edge, seed[i], kTagMain);
                X[i] = -p0 * log_tab(u01_open0(w.x, w.y), lt);
            }
            break;
        case kExpTrunc: {
            // Inversion of the CDF truncated at m: U' = q + (1 - q) U is uniform on
            // (q, 1] with q = e^{-m/mean}, and X = -mean log U' lies in [0, m].
            const DistParams D = dists[(info >> kInfoDistShift) & kInfoDistMask];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const Words2 w = draw_words2(edge, seed[i], kTagMain);
                const double u = __builtin_fma(u01_open0(w.x, w.y), D.p[3], D.p[1]);
                const double x = -p0 * log_tab(u, lt);
                X[i] = (x > D.p[2]) ? D.p[2] : x;
            }
            break;
        }
        default: {
            if (!(info & kInfoTruncated) && kind == kGammaErlang2) {  // sum of two unit exponentials
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    const Words4 w = draw_words4(edge, seed[i], 0u, kTagMain);
                    X[i] = -p0 * log_tab(u01_open0(w.x, w.y) * u01_open0(w.z, w.w), lt);
                }
            } else if (!(info & kInfoTruncated) && kind == kGammaHalf) {  // Exp(1) cos^2(2 pi U)
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    const Words4 w = draw_words4(edge, seed[i], 0u, kTagMain);
                    const double cs = cos2pi_fast(u01_mid32(w.z));
                    X[i] = p0 * (-log_tab(u01_open0(w.x, w.y), lt) * (cs * cs));
                }
            } else {
                const DistParams D = dists[(info >> kInfoDistShift) & kInfoDistMask];
#pragma unroll
                for (int i = 0; i < N; ++i) X[i] = gamma_rejection(kind, info, D, edge, seed[i], lt);
            }
            break;
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double extra = X[i] * dmin;
        dur[i] = dmin + extra;
    }
}

}  // namespace mcdp
