// gfx950 kernels for per-event statistics over event-major results
// (realized[e * ld + s], cause[e * ld + s], s < n): one workgroup per event,
// reading its row of n samples.
//   * mean: the sum of the row, exactly rounded (double-double accumulation,
//     a fixed reduction order), divided by n;
//   * quantiles: numpy's default method ("linear", Hyndman & Fan 7) from
//     exact order statistics found by radix selection on order-preserving
//     64-bit keys (8 passes of 8 bits; the row stays in L2 between passes);
//   * cause counts: how many samples have cause -1 and how many have each
//     distinct predecessor source, counted in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcdp_stats.hpp"

namespace mcdp {

namespace {

constexpr int kBlock = 256;

// s + e == a + b exactly, s = fl(a + b) (Knuth's TwoSum; this file is
// compiled without FMA contraction).
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD x, double y) {
    double s, e;
    two_sum(x.hi, y, s, e);
    e += x.lo;
    DD r;
    two_sum(s, e, r.hi, r.lo);
    return r;
}

__device__ __forceinline__ DD dd_add(DD x, DD y) {
    double s, e;
    two_sum(x.hi, y.hi, s, e);
    e += x.lo + y.lo;
    DD r;
    two_sum(s, e, r.hi, r.lo);
    return r;
}

// Order-preserving key of a double: -inf < ... < -0 < +0 < ... < +inf (NaNs
// are counted apart and never selected).
__device__ __forceinline__ uint64_t order_key(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_value(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

__global__ void __launch_bounds__(kBlock) mean_kernel(const double *__restrict__ realized, int32_t n, uint32_t ld,
                                                      int32_t E, double *__restrict__ mean) {
    __shared__ DD part[kBlock];
    __shared__ double naive[kBlock];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const double *row = realized + (uint64_t)(uint32_t)e * ld;
        DD acc{0.0, 0.0};
        double plain = 0.0;
        for (int32_t s = threadIdx.x; s < n; s += kBlock) {
            const double v = row[s];
            acc = dd_add(acc, v);
            plain += v;
        }
        part[threadIdx.x] = acc;
        naive[threadIdx.x] = plain;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {  // fixed pairing: deterministic
            if ((int)threadIdx.x < w) {
                part[threadIdx.x] = dd_add(part[threadIdx.x], part[threadIdx.x + w]);
                naive[threadIdx.x] += naive[threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            // A non-finite sum (inf, NaN) carries no rounding to correct.
            const double total = __builtin_isfinite(naive[0]) ? part[0].hi : naive[0];
            mean[e] = total / (double)n;
        }
        __syncthreads();
    }
}

// numpy.quantile(row, q, method="linear") for each q, given the row's order
// statistics: virtual index (n - 1) q, neighbours floor / floor + 1 (clamped
// to [0, n - 1]), gamma = virtual - floor, and numpy's two-sided lerp.
__device__ double lerp_numpy(double a, double b, double t) {
    const double diff = b - a;
    return (t >= 0.5) ? b - diff * (1.0 - t) : a + diff * t;
}

__global__ void __launch_bounds__(kBlock) quantile_kernel(const double *__restrict__ realized, int32_t n, uint32_t ld,
                                                          int32_t E, QuantileArgs qa, double *__restrict__ out) {
    __shared__ uint32_t hist[kMaxRanks][256];
    __shared__ uint64_t prefix[kMaxRanks];
    __shared__ int32_t krem[kMaxRanks];
    __shared__ uint32_t nans;
    const int R = qa.n_ranks;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const double *row = realized + (uint64_t)(uint32_t)e * ld;
        if ((int)threadIdx.x < R) {
            prefix[threadIdx.x] = 0;
            krem[threadIdx.x] = qa.rank[threadIdx.x];
        }
        if (threadIdx.x == 0) nans = 0;
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            for (int i = threadIdx.x; i < R * 256; i += kBlock) hist[i >> 8][i & 255] = 0;
            __syncthreads();
            // digits of each rank's key found so far (above this pass's digit)
            uint64_t pre[kMaxRanks];
#pragma unroll
            for (int r = 0; r < kMaxRanks; ++r) pre[r] = (r < R && pass > 0) ? prefix[r] >> (shift + 8) : 0;
            uint32_t my_nans = 0;
            const int lane = threadIdx.x & 63;
            for (int32_t s = threadIdx.x; s < n; s += kBlock) {
                const double v = row[s];
                const bool ok = v == v;
                my_nans += (!ok && pass == 0) ? 1u : 0u;
                const uint64_t key = order_key(v);
                const uint32_t digit = (uint32_t)(key >> shift) & 255u;
                const uint64_t high = pass > 0 ? key >> (shift + 8) : 0;
                const uint64_t active = __ballot(1);
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(digit);
                const bool same_digit = __all(digit == d0);
#pragma unroll
                for (int r = 0; r < kMaxRanks; ++r) {
                    if (r < R) {
                        const bool m = ok && high == pre[r];
                        const uint64_t mm = __ballot(m);
                        if (mm == active && same_digit) {
                            // the whole wave adds to one bin (rows saturated at a
                            // cap hold many equal values): one atomic, not 64
                            if (lane == __ffsll((long long)active) - 1)
                                atomicAdd(&hist[r][d0], (uint32_t)__popcll(mm));
                        } else if (m) {
                            atomicAdd(&hist[r][digit], 1u);
                        }
                    }
                }
            }
            if (pass == 0 && my_nans) atomicAdd(&nans, my_nans);
            __syncthreads();
            if ((int)threadIdx.x < R) {
                const int r = threadIdx.x;
                int32_t k = krem[r];
                uint32_t d = 0;
                for (; d < 255; ++d) {
                    const int32_t c = (int32_t)hist[r][d];
                    if (k < c) break;
                    k -= c;
                }
                krem[r] = k;
                prefix[r] |= (uint64_t)d << shift;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            for (int j = 0; j < qa.n_q; ++j) {
                double res;
                if (nans) {
                    res = __builtin_nan("");
                } else {
                    const double a = key_value(prefix[qa.lo[j]]);
                    const double b = key_value(prefix[qa.hi[j]]);
                    res = lerp_numpy(a, b, qa.gamma[j]);
                }
                out[(uint64_t)(uint32_t)e * (uint32_t)qa.n_q + j] = res;
            }
        }
        __syncthreads();
    }
}

// Index of `c` among the event's distinct sources (ascending keys), or -1.
__device__ __forceinline__ int32_t find_source(const int32_t *__restrict__ keys, int32_t nk, int32_t c) {
    int32_t lo = 0, hi = nk;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (keys[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < nk && keys[lo] == c) ? lo : -1;
}

__global__ void __launch_bounds__(kBlock) cause_kernel(const int32_t *__restrict__ cause, int32_t n, uint32_t ld,
                                                       int32_t E, const int32_t *__restrict__ pred_off,
                                                       const int32_t *__restrict__ pred_key,
                                                       const int32_t *__restrict__ pred_pos,
                                                       const int32_t *__restrict__ pred_nkey,
                                                       int64_t *__restrict__ counts) {
    __shared__ uint32_t c_lds[kCauseLdsCounters];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const int32_t p0 = pred_off[e], np = pred_off[e + 1] - p0, nk = pred_nkey[e];
        const int32_t *keys = pred_key + p0;
        const int32_t *pos = pred_pos + p0;
        int64_t *out = counts + (int64_t)p0 + e;  // [1 + np]: cause -1, then each predecessor position
        const bool in_lds = np + 1 <= kCauseLdsCounters;
        if (in_lds) {
            for (int i = threadIdx.x; i <= np; i += kBlock) c_lds[i] = 0;
        } else {
            for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = 0;
        }
        __syncthreads();
        const int32_t *row = cause + (uint64_t)(uint32_t)e * ld;
        for (int32_t s = threadIdx.x; s < n; s += kBlock) {
            const int32_t c = row[s];
            int32_t slot = -1;
            if (c == -1) {
                slot = 0;
            } else {
                const int32_t k = find_source(keys, nk, c);
                if (k >= 0) slot = 1 + pos[k];
            }
            if (slot < 0) continue;  // not a predecessor of e: counted nowhere
            if (in_lds)
                atomicAdd(&c_lds[slot], 1u);
            else
                atomicAdd((unsigned long long *)&out[slot], 1ull);
        }
        __syncthreads();
        if (in_lds)
            for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = (int64_t)c_lds[i];
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_event_mean(const double *realized, int32_t n, int32_t ld, int32_t E, double *mean,
                             hipStream_t stream) {
    if (E > 0) hipLaunchKernelGGL(mean_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, realized, n, (uint32_t)ld, E, mean);
    return hipGetLastError();
}

hipError_t launch_event_quantiles(const double *realized, int32_t n, int32_t ld, int32_t E, const QuantileArgs &qa,
                                  double *out, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(quantile_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, realized, n, (uint32_t)ld, E, qa,
                           out);
    return hipGetLastError();
}

hipError_t launch_cause_counts(const int32_t *cause, int32_t n, int32_t ld, int32_t E, const int32_t *pred_off,
                               const int32_t *pred_key, const int32_t *pred_pos, const int32_t *pred_nkey,
                               int64_t *counts, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(cause_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, cause, n, (uint32_t)ld, E, pred_off,
                           pred_key, pred_pos, pred_nkey, counts);
    return hipGetLastError();
}

}  // namespace mcdp
