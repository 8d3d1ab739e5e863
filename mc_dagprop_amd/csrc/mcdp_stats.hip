// gfx950 kernels for per-event statistics over event-major results
// (realized[e * ld + s], cause[e * ld + s], s < n): one workgroup per event,
// reading its row of n samples.
//   * mean: the sum of the row, exactly rounded (double-double accumulation,
//     a fixed reduction order), divided by n;
//   * quantiles: numpy's default method ("linear", Hyndman & Fan 7) from
//     exact order statistics: a first pass finds the row's extremes and their
//     multiplicities (ranks inside those masses are done), then radix
//     selection on order-preserving 64-bit keys over the key bytes in which
//     the extremes differ (at most 8 passes of 8 bits);
//   * cause counts: how many samples have cause -1 and how many have each
//     distinct predecessor source, counted in registers (events with at most
//     kCauseRegKeys distinct sources) or in LDS.
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

constexpr int kWaves = kBlock / 64;
static_assert(kWaves == 4, "block reductions merge four waves");

// Butterfly sum over a wave's 64 lanes in a fixed order: the double-double
// sum and the plain one.  Every step is commutative (TwoSum's error term is
// exact, so a pair's lanes compute the same double-double), so all lanes end
// with the same value.
__device__ __forceinline__ void wave_sum(DD &acc, double &plain) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        acc = dd_add(acc, DD{__shfl_xor(acc.hi, o), __shfl_xor(acc.lo, o)});
        plain += __shfl_xor(plain, o);
    }
}

// Order-preserving key of a double: -inf < ... < -0 < +0 < ... < +inf (NaN
// keys lie beyond the infinities' and are never selected).
__device__ __forceinline__ uint64_t order_key(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_value(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// The mean of one row by a block of kBlock threads (part, naive: LDS of
// kWaves entries each).
__device__ __forceinline__ void mean_row(const double *__restrict__ row, int32_t n, double *__restrict__ mean,
                                         DD *part, double *naive) {
    {
        DD acc{0.0, 0.0};
        double plain = 0.0;
        // four loads per lane in flight; samples tid, tid + kBlock, ... in order
        auto visit = [&](double v) {
            acc = dd_add(acc, v);
            plain += v;
        };
        int32_t s = threadIdx.x;
        for (; s + 3 * kBlock < n; s += 4 * kBlock) {
            const double v0 = row[s], v1 = row[s + kBlock], v2 = row[s + 2 * kBlock], v3 = row[s + 3 * kBlock];
            visit(v0);
            visit(v1);
            visit(v2);
            visit(v3);
        }
        for (; s < n; s += kBlock) visit(row[s]);
        wave_sum(acc, plain);
        if ((threadIdx.x & 63) == 0) {
            part[threadIdx.x >> 6] = acc;
            naive[threadIdx.x >> 6] = plain;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // the waves' sums in a fixed order; a non-finite sum (inf, NaN)
            // carries no rounding to correct
            const DD sum = dd_add(dd_add(part[0], part[1]), dd_add(part[2], part[3]));
            const double nv = (naive[0] + naive[1]) + (naive[2] + naive[3]);
            *mean = (__builtin_isfinite(nv) ? sum.hi : nv) / (double)n;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) mean_kernel(const double *__restrict__ realized, int32_t n, uint32_t ld,
                                                      int32_t E, double *__restrict__ mean) {
    __shared__ DD part[kWaves];
    __shared__ double naive[kWaves];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x)
        mean_row(realized + (uint64_t)(uint32_t)e * ld, n, mean + e, part, naive);
}

// numpy.quantile(row, q, method="linear") for each q, given the row's order
// statistics: virtual index (n - 1) q, neighbours floor / floor + 1 (clamped
// to [0, n - 1]), gamma = virtual - floor, and numpy's two-sided lerp.
__device__ double lerp_numpy(double a, double b, double t) {
    const double diff = b - a;
    return (t >= 0.5) ? b - diff * (1.0 - t) : a + diff * t;
}

// Adds this wave's matching lanes to hist[digit]: lanes that share the first
// matching lane's digit (the common case for rows with many equal values, e.g.
// saturated at a cap) in one atomic, the others one atomic each.
__device__ __forceinline__ void wave_hist_add(uint32_t *hist, bool m, uint32_t digit, int lane) {
    const uint64_t mm = __ballot(m);
    if (mm == 0) return;
    const int leader = __ffsll((long long)mm) - 1;
    const uint32_t d0 = __shfl(digit, leader);
    const uint64_t same = __ballot(m && digit == d0);
    if (lane == leader) atomicAdd(&hist[d0], (uint32_t)__popcll(same));
    if (m && digit != d0) atomicAdd(&hist[digit], 1u);
}

// Extremes of a row for the quantile kernel: smallest and largest key with
// their multiplicities, and the NaN count.
struct Extremes {
    uint64_t mn, mx;
    uint32_t cmn, cmx, nans;
};

__device__ __forceinline__ Extremes merge_extremes(Extremes a, const Extremes &b) {
    if (b.mn < a.mn) {
        a.mn = b.mn;
        a.cmn = b.cmn;
    } else if (b.mn == a.mn) {
        a.cmn += b.cmn;
    }
    if (b.mx > a.mx) {
        a.mx = b.mx;
        a.cmx = b.cmx;
    } else if (b.mx == a.mx) {
        a.cmx += b.cmx;
    }
    a.nans += b.nans;
    return a;
}

// Butterfly merge of the extremes over a wave's 64 lanes (as wave_sum).
__device__ __forceinline__ Extremes wave_merge_extremes(Extremes x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const Extremes y{__shfl_xor(x.mn, o), __shfl_xor(x.mx, o), __shfl_xor(x.cmn, o), __shfl_xor(x.cmx, o),
                         __shfl_xor(x.nans, o)};
        x = merge_extremes(x, y);
    }
    return x;
}

// The quantile kernel's workgroup: rows whose ranks need passes beyond the
// first are streamed again by one workgroup while the others run, so a
// larger one keeps more of such a row in flight.
constexpr int kQBlock = 512;
constexpr int kQWaves = kQBlock / 64;

// The quantile kernel's first pass over a row: this thread's share of the
// extremes with their multiplicities, of the plain sum and, with kSum, of the
// double-double sum.  Extremes are taken over values: NaNs are left out (a
// row holding one has a NaN plain sum, which the caller then checks), and -0
// and +0 count as one value, keyed as -0 for a smallest value and as +0 for a
// largest one, so that the key range brackets both zeros.  Four loads per
// lane are in flight per iteration, and the four values are reduced to their
// smallest and largest before they meet the running extremes.  Each thread
// visits its samples tid, tid + kB, ... in that order, so the sums'
// rounding does not depend on the unrolling.
template <int kB, bool kSum>
__device__ __forceinline__ void extremes_pass(const double *__restrict__ row, int32_t n, Extremes &x, DD &acc,
                                              double &plain) {
    double mn = __builtin_inf(), mx = -__builtin_inf();
    uint32_t cmn = 0, cmx = 0;
    auto sum = [&](double v) {
        if (kSum) acc = dd_add(acc, v);
        plain += v;
    };
    int32_t s = threadIdx.x;
    for (; s + 3 * kB < n; s += 4 * kB) {
        const double v0 = row[s], v1 = row[s + kB], v2 = row[s + 2 * kB], v3 = row[s + 3 * kB];
        sum(v0);
        sum(v1);
        sum(v2);
        sum(v3);
        // fmin / fmax pass over a NaN operand; four NaNs give a NaN, which the
        // compares below never take
        const double lo = __builtin_fmin(__builtin_fmin(v0, v1), __builtin_fmin(v2, v3));
        const double hi = __builtin_fmax(__builtin_fmax(v0, v1), __builtin_fmax(v2, v3));
        const bool lt = lo < mn;
        mn = lt ? lo : mn;
        cmn = (lt ? 0u : cmn) + (v0 == mn ? 1u : 0u) + (v1 == mn ? 1u : 0u) + (v2 == mn ? 1u : 0u) +
              (v3 == mn ? 1u : 0u);
        const bool gt = hi > mx;
        mx = gt ? hi : mx;
        cmx = (gt ? 0u : cmx) + (v0 == mx ? 1u : 0u) + (v1 == mx ? 1u : 0u) + (v2 == mx ? 1u : 0u) +
              (v3 == mx ? 1u : 0u);
    }
    for (; s < n; s += kB) {
        const double v = row[s];
        sum(v);
        const bool lt = v < mn;
        mn = lt ? v : mn;
        cmn = (lt ? 0u : cmn) + (v == mn ? 1u : 0u);
        const bool gt = v > mx;
        mx = gt ? v : mx;
        cmx = (gt ? 0u : cmx) + (v == mx ? 1u : 0u);
    }
    x.mn = order_key(mn == 0.0 ? -0.0 : mn);
    x.mx = order_key(mx == 0.0 ? 0.0 : mx);
    x.cmn = cmn;
    x.cmx = cmx;
    x.nans = 0;
}

constexpr int kValueBins = kMaxRanks * 256;  // the radix histograms' LDS, reused
constexpr int kGather = kValueBins / 2;      // keys of one bin gathered into the same LDS

// Rows with ranks outside the masses of their extremes, before any radix
// pass: one pass counts the row in kValueBins bins of equal width over
// [smallest, largest] value, bin floor((v - smallest) * kValueBins / range).
// The bin is monotone in v, so the k-th smallest sample lies in the bin where
// the running count passes k.  A second pass per distinct bin gathers that
// bin's keys into LDS (at most kGather), where each rank in it is selected
// exactly by counting, for every gathered key, the keys below and equal to it.
// Ranks in bins holding more than kGather keys, and rows whose range is not a
// positive finite double with a finite bin scale, are left to the radix
// passes (done[r] stays 0).  Rows with NaNs never get here (every rank done).
template <int kB>
__device__ void value_bin_select(const double *__restrict__ row, int32_t n, const Extremes &X, int R, int lane,
                                 uint32_t *vh, uint32_t *seg, uint64_t *prefix, const int32_t *krem, int32_t *done,
                                 int32_t *vbin, int32_t *vloc, int32_t *vcnt, uint32_t &cand_n) {
    bool open = false;
    for (int r = 0; r < R; ++r) open |= done[r] == 0;
    const double lo = key_value(X.mn), hi = key_value(X.mx);
    const double span = hi - lo;
    const double scale = (double)kValueBins / span;
    if (!open || !(span > 0.0 && span < __builtin_inf() && scale < __builtin_inf()))
        return;  // uniform: LDS and X agree in all threads
    auto bin_of = [&](double v) {
        const double f = (v - lo) * scale;
        return f < (double)kValueBins ? (int32_t)f : kValueBins - 1;
    };
    // four loads per lane in flight
    auto scan = [&](auto &&visit) {
        int32_t s = threadIdx.x;
        for (; s + 3 * kB < n; s += 4 * kB) {
            double v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = row[s + i * kB];
#pragma unroll
            for (int i = 0; i < 4; ++i) visit(v[i]);
        }
        for (; s < n; s += kB) visit(row[s]);
    };
    for (int i = threadIdx.x; i < kValueBins; i += kB) vh[i] = 0;
    __syncthreads();
    scan([&](double v) { wave_hist_add(vh, true, (uint32_t)bin_of(v), lane); });
    __syncthreads();
    static_assert(kValueBins % kB == 0, "value bins split evenly over the block");
    constexpr int per = kValueBins / kB;
    uint32_t t = 0;
    for (int i = 0; i < per; ++i) t += vh[threadIdx.x * per + i];
    seg[threadIdx.x] = t;
    __syncthreads();
    if ((int)threadIdx.x < R && !done[threadIdx.x]) {
        const int r = threadIdx.x;
        int32_t k = krem[r];  // its rank in the row (k < n)
        int g = 0;
        while (k >= (int32_t)seg[g]) k -= (int32_t)seg[g++];
        int b = g * per;
        while (k >= (int32_t)vh[b]) k -= (int32_t)vh[b++];
        vbin[r] = b;
        vloc[r] = k;
        vcnt[r] = (int32_t)vh[b];
    }
    __syncthreads();
    uint64_t *cand = reinterpret_cast<uint64_t *>(vh);
    for (int r = 0; r < R; ++r) {
        if (done[r] || vcnt[r] > kGather) continue;  // uniform (LDS, after a barrier)
        const int32_t b = vbin[r];
        if (threadIdx.x == 0) cand_n = 0;
        __syncthreads();
        scan([&](double v) {
            const bool m = bin_of(v) == b;
            const uint64_t mm = __ballot(m);
            if (mm == 0) return;
            const int leader = __ffsll((long long)mm) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&cand_n, (uint32_t)__popcll(mm));
            base = __shfl(base, leader);
            if (m) cand[base + (uint32_t)__popcll(mm & ((1ull << lane) - 1ull))] = order_key(v);
        });
        __syncthreads();
        const int32_t c = (int32_t)cand_n;  // == vcnt[r]
        for (int32_t i = threadIdx.x; i < c; i += kB) {
            const uint64_t ki = cand[i];
            int32_t below = 0, same = 0;
            for (int32_t j = 0; j < c; ++j) {
                const uint64_t kj = cand[j];
                below += kj < ki ? 1 : 0;
                same += kj == ki ? 1 : 0;
            }
            for (int q = 0; q < R; ++q)  // equal keys write the same value
                if (!done[q] && vbin[q] == b && below <= vloc[q] && vloc[q] < below + same) prefix[q] = ki;
        }
        __syncthreads();
        if ((int)threadIdx.x < R && !done[threadIdx.x] && vbin[threadIdx.x] == b) done[threadIdx.x] = 1;
        __syncthreads();
    }
}

// LDS of the quantile kernels' work on one row.
struct QuantileLds {
    uint32_t hist[kMaxRanks][256];
    Extremes ext[kQWaves];
    DD part[kQWaves];
    double naive[kQWaves];
    uint64_t prefix[kMaxRanks];   // per rank: its key's digits found so far
    int32_t krem[kMaxRanks];      // per rank: its rank among keys with that prefix
    int32_t done[kMaxRanks];      // per rank: key known
    uint64_t gprefix[kMaxRanks];  // ranks whose prefixes agree share one histogram
    int32_t group_of[kMaxRanks];
    int32_t n_groups;
    int32_t first_pass;
    uint32_t seg[kQBlock];        // value bins: counts per run of kValueBins / kQBlock bins
    int32_t vbin[kMaxRanks];      // per rank: its value bin, rank inside it, the bin's count
    int32_t vloc[kMaxRanks];
    int32_t vcnt[kMaxRanks];
    uint32_t cand_n;
};

// The quantiles of one row by a block of kQBlock threads, out[0 .. n_q); with
// kMean, its first pass also sums the row into *mean (the mean without a
// separate read).
template <bool kMean>
__device__ __forceinline__ void quantile_row(const double *__restrict__ row, int32_t n, const QuantileArgs &qa,
                                             double *__restrict__ out, double *__restrict__ mean, QuantileLds &L) {
    auto &hist = L.hist;
    auto &ext = L.ext;
    auto &part = L.part;
    auto &naive = L.naive;
    auto &prefix = L.prefix;
    auto &krem = L.krem;
    auto &done = L.done;
    auto &gprefix = L.gprefix;
    auto &group_of = L.group_of;
    auto &n_groups = L.n_groups;
    auto &first_pass = L.first_pass;
    auto &seg = L.seg;
    auto &vbin = L.vbin;
    auto &vloc = L.vloc;
    auto &vcnt = L.vcnt;
    auto &cand_n = L.cand_n;
    const int R = qa.n_ranks;
    const int lane = threadIdx.x & 63;
    {
        // Pass 0: extremes.  Ranks inside the mass of the smallest or the
        // largest value (a row saturated at its cap puts its high quantiles
        // there) are known at once; the others need only the key bytes in
        // which the smallest and largest keys differ.
        Extremes x{~0ull, 0ull, 0u, 0u, 0u};
        DD acc{0.0, 0.0};
        double plain = 0.0;
        extremes_pass<kQBlock, kMean>(row, n, x, acc, plain);
        x = wave_merge_extremes(x);
        if (kMean) {
            wave_sum(acc, plain);
        } else {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) plain += __shfl_xor(plain, o);
        }
        if (lane == 0) {
            ext[threadIdx.x >> 6] = x;
            part[threadIdx.x >> 6] = acc;
            naive[threadIdx.x >> 6] = plain;
        }
        __syncthreads();
        // the waves' results in wave order
        Extremes X = ext[0];
        double nv = naive[0];
        for (int w = 1; w < kQWaves; ++w) {
            X = merge_extremes(X, ext[w]);
            nv += naive[w];
        }
        if (kMean && threadIdx.x == 0) {
            DD sum = part[0];
            for (int w = 1; w < kQWaves; ++w) sum = dd_add(sum, part[w]);
            *mean = (__builtin_isfinite(nv) ? sum.hi : nv) / (double)n;
        }
        // A NaN sample makes the plain sum NaN, and so does +inf + -inf: only
        // then is the row read again for NaNs (nv is the same in every thread).
        bool has_nan = false;
        if (nv != nv) {
            int found = 0;
            for (int32_t s = threadIdx.x; s < n; s += kQBlock) found |= row[s] != row[s];
            has_nan = __syncthreads_or(found) != 0;
        }
        if ((int)threadIdx.x < R) {
            const int r = threadIdx.x;
            const int32_t k = qa.rank[r];
            done[r] = 1;
            if (has_nan || k >= n - (int32_t)X.cmx) {
                prefix[r] = X.mx;
            } else if (k < (int32_t)X.cmn) {
                prefix[r] = X.mn;
            } else {
                done[r] = 0;
                prefix[r] = 0;
                krem[r] = k;
            }
        }
        if (threadIdx.x == 0) {
            // whole key bytes shared by every key of the row (mn < mx here if
            // any rank is left)
            const uint64_t diff = X.mn ^ X.mx;
            first_pass = diff ? __builtin_clzll(diff) / 8 : 8;
        }
        __syncthreads();
        const int p0 = first_pass;
        if ((int)threadIdx.x < R && !done[threadIdx.x] && p0 > 0)
            prefix[threadIdx.x] = X.mn & ~(~0ull >> (8 * p0));
        __syncthreads();
        value_bin_select<kQBlock>(row, n, X, R, lane, &hist[0][0], seg, prefix, krem, done, vbin, vloc, vcnt, cand_n);
        for (int pass = p0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            __syncthreads();
            if (threadIdx.x == 0) {  // distinct prefixes among the open ranks
                int g = 0;
                for (int r = 0; r < R; ++r) {
                    group_of[r] = -1;
                    if (done[r]) continue;
                    int found = -1;
                    for (int h = 0; h < g; ++h)
                        if (gprefix[h] == prefix[r]) found = h;
                    if (found < 0) {
                        gprefix[g] = prefix[r];
                        found = g++;
                    }
                    group_of[r] = found;
                }
                n_groups = g;
            }
            __syncthreads();
            const int G = n_groups;
            if (G == 0) break;
            for (int i = threadIdx.x; i < G * 256; i += kQBlock) hist[i >> 8][i & 255] = 0;
            __syncthreads();
            auto visit = [&](double v) {
                const bool ok = v == v;
                const uint64_t key = order_key(v);
                const uint32_t digit = (uint32_t)(key >> shift) & 255u;
                const uint64_t high = pass > 0 ? key >> (shift + 8) : 0;
                for (int g = 0; g < G; ++g)  // the groups' prefixes from LDS (broadcast reads)
                    wave_hist_add(hist[g], ok && high == (pass > 0 ? gprefix[g] >> (shift + 8) : 0), digit, lane);
            };
            // Few rows need these passes, and each runs in one workgroup while
            // the others have finished: eight loads per lane in flight keep
            // such a row from being bound by memory latency.
            int32_t s = threadIdx.x;
            for (; s + 7 * kQBlock < n; s += 8 * kQBlock) {
                double v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = row[s + i * kQBlock];
#pragma unroll
                for (int i = 0; i < 8; ++i) visit(v[i]);
            }
            for (; s < n; s += kQBlock) visit(row[s]);
            __syncthreads();
            if ((int)threadIdx.x < R && !done[threadIdx.x]) {
                const int r = threadIdx.x;
                const uint32_t *h = hist[group_of[r]];
                int32_t k = krem[r];
                uint32_t d = 0;
                for (; d < 255; ++d) {
                    const int32_t c = (int32_t)h[d];
                    if (k < c) break;
                    k -= c;
                }
                krem[r] = k;
                prefix[r] |= (uint64_t)d << shift;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = 0; j < qa.n_q; ++j) {
                double res;
                if (has_nan) {
                    res = __builtin_nan("");
                } else {
                    const double a = key_value(prefix[qa.lo[j]]);
                    const double b = key_value(prefix[qa.hi[j]]);
                    res = lerp_numpy(a, b, qa.gamma[j]);
                }
                out[j] = res;
            }
        }
        __syncthreads();  // prefix[] and ext[] are read above before the next row resets them
    }
}

// Quantiles of each row; with kMean, also the row means.
template <bool kMean>
__global__ void __launch_bounds__(kQBlock) quantile_kernel(const double *__restrict__ realized, int32_t n, uint32_t ld,
                                                          int32_t E, QuantileArgs qa, double *__restrict__ out,
                                                          double *__restrict__ mean) {
    __shared__ QuantileLds L;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x)
        quantile_row<kMean>(realized + (uint64_t)(uint32_t)e * ld, n, qa, out + (uint64_t)(uint32_t)e * (uint32_t)qa.n_q,
                            kMean ? mean + e : nullptr, L);
}

// Events with at most this many distinct sources count their row in registers.
constexpr int kCauseRegKeys = 8;

// First sample of a row of causes at a 16-byte boundary (at most n).
__device__ __forceinline__ int32_t aligned_head(const int32_t *row, int32_t n) {
    return min(n, (int32_t)(((16u - ((uint32_t)(uintptr_t)row & 15u)) & 15u) / 4u));
}

// Cause counts of one row for an event with NK distinct sources: per-lane
// counters in registers (cnt[0] for cause -1, cnt[1 + k] for the source
// keys[k]), one compare and add per counter and sample, then one LDS atomic
// per counter and wave into c_lds[0] and c_lds[1 + pos[k]].  Causes that are
// no predecessor of the event are counted nowhere.  Two 16-byte loads per
// lane are in flight per iteration over the row's aligned middle.
template <int NK>
__device__ __forceinline__ void count_row_regs(const int32_t *__restrict__ row, int32_t n,
                                               const int32_t *__restrict__ keys, const int32_t *__restrict__ pos,
                                               uint32_t *c_lds) {
    int32_t key[NK > 0 ? NK : 1];
#pragma unroll
    for (int k = 0; k < NK; ++k) key[k] = keys[k];
    uint32_t cnt[NK + 1];
#pragma unroll
    for (int k = 0; k <= NK; ++k) cnt[k] = 0;
    auto count = [&](int32_t c) {
        cnt[0] += c == -1 ? 1u : 0u;
#pragma unroll
        for (int k = 0; k < NK; ++k) cnt[1 + k] += c == key[k] ? 1u : 0u;
    };
    const int32_t head = aligned_head(row, n);
    const int32_t n_vec = (n - head) / 4;
    const int4 *vrow = reinterpret_cast<const int4 *>(row + head);
    for (int32_t s = threadIdx.x; s < head; s += kBlock) count(row[s]);
    int32_t v = threadIdx.x;
    for (; v + kBlock < n_vec; v += 2 * kBlock) {
        const int4 a = vrow[v], b = vrow[v + kBlock];
        count(a.x);
        count(a.y);
        count(a.z);
        count(a.w);
        count(b.x);
        count(b.y);
        count(b.z);
        count(b.w);
    }
    for (; v < n_vec; v += kBlock) {
        const int4 a = vrow[v];
        count(a.x);
        count(a.y);
        count(a.z);
        count(a.w);
    }
    for (int32_t s = head + 4 * n_vec + (int32_t)threadIdx.x; s < n; s += kBlock) count(row[s]);
#pragma unroll
    for (int k = 0; k <= NK; ++k) {
        uint32_t t = cnt[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(&c_lds[k == 0 ? 0 : 1 + pos[k - 1]], t);
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

// LDS of the cause kernels' work on one row.
struct CauseLds {
    uint32_t c[kCauseLdsCounters];
    int32_t k[kCauseLdsKeys], p[kCauseLdsKeys];
};

// The cause counts of event e's row (n samples) by a block of kBlock threads.
__device__ __forceinline__ void cause_row(const int32_t *__restrict__ row, int32_t n, int32_t e,
                                          const int32_t *__restrict__ pred_off, const int32_t *__restrict__ pred_key,
                                          const int32_t *__restrict__ pred_pos,
                                          const int32_t *__restrict__ pred_nkey, int64_t *__restrict__ counts,
                                          int accumulate, CauseLds &L) {
    auto &c_lds = L.c;
    auto &k_lds = L.k;
    auto &p_lds = L.p;
    {
        const int32_t p0 = pred_off[e], np = pred_off[e + 1] - p0, nk = pred_nkey[e];
        // the event's distinct sources and positions, in LDS when they fit
        const bool keys_in_lds = nk <= kCauseLdsKeys;
        for (int i = threadIdx.x; keys_in_lds && i < nk; i += kBlock) {
            k_lds[i] = pred_key[p0 + i];
            p_lds[i] = pred_pos[p0 + i];
        }
        const int32_t *keys = keys_in_lds ? k_lds : pred_key + p0;
        const int32_t *pos = keys_in_lds ? p_lds : pred_pos + p0;
        int64_t *out = counts + (int64_t)p0 + e;  // [1 + np]: cause -1, then each predecessor position
        const bool in_lds = np + 1 <= kCauseLdsCounters;
        if (in_lds) {
            for (int i = threadIdx.x; i <= np; i += kBlock) c_lds[i] = 0;
        } else if (!accumulate) {
            for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = 0;
        }
        __syncthreads();
        if (in_lds && nk <= kCauseRegKeys) {
            const int32_t *k0 = pred_key + p0, *q0 = pred_pos + p0;
            switch (nk) {
            case 0: count_row_regs<0>(row, n, k0, q0, c_lds); break;
            case 1: count_row_regs<1>(row, n, k0, q0, c_lds); break;
            case 2: count_row_regs<2>(row, n, k0, q0, c_lds); break;
            case 3: count_row_regs<3>(row, n, k0, q0, c_lds); break;
            case 4: count_row_regs<4>(row, n, k0, q0, c_lds); break;
            case 5: count_row_regs<5>(row, n, k0, q0, c_lds); break;
            case 6: count_row_regs<6>(row, n, k0, q0, c_lds); break;
            case 7: count_row_regs<7>(row, n, k0, q0, c_lds); break;
            default: count_row_regs<8>(row, n, k0, q0, c_lds); break;
            }
            __syncthreads();
            for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = (accumulate ? out[i] : 0) + (int64_t)c_lds[i];
            __syncthreads();
            return;
        }
        auto count = [&](int32_t c) {
            int32_t slot = -1;
            if (c == -1) {
                slot = 0;
            } else {
                const int32_t k = find_source(keys, nk, c);
                if (k >= 0) slot = 1 + pos[k];
            }
            // slot < 0: not a predecessor of e, counted nowhere
            if (in_lds) {
                // a wave's causes take few distinct values: one atomic per value
                bool pending = slot >= 0;
                while (__ballot(pending)) {
                    const uint64_t mm = __ballot(pending);
                    const int leader = __ffsll((long long)mm) - 1;
                    const int32_t s0 = __shfl(slot, leader);
                    const uint64_t same = __ballot(pending && slot == s0);
                    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&c_lds[s0], (uint32_t)__popcll(same));
                    pending = pending && slot != s0;
                }
            } else if (slot >= 0) {
                atomicAdd((unsigned long long *)&out[slot], 1ull);
            }
        };
        // 16-byte loads over the row's aligned middle (four samples a lane),
        // single samples before and after it
        const int32_t head = aligned_head(row, n);
        const int32_t n_vec = (n - head) / 4;
        const int4 *vrow = reinterpret_cast<const int4 *>(row + head);
        for (int32_t s = threadIdx.x; s < head; s += kBlock) count(row[s]);
        for (int32_t v = threadIdx.x; v < n_vec; v += kBlock) {
            const int4 c4 = vrow[v];
            count(c4.x);
            count(c4.y);
            count(c4.z);
            count(c4.w);
        }
        for (int32_t s = head + 4 * n_vec + (int32_t)threadIdx.x; s < n; s += kBlock) count(row[s]);
        __syncthreads();
        if (in_lds)
            for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = (accumulate ? out[i] : 0) + (int64_t)c_lds[i];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) cause_kernel(const int32_t *__restrict__ cause, int32_t n, uint32_t ld,
                                                       int32_t E, const int32_t *__restrict__ pred_off,
                                                       const int32_t *__restrict__ pred_key,
                                                       const int32_t *__restrict__ pred_pos,
                                                       const int32_t *__restrict__ pred_nkey,
                                                       int64_t *__restrict__ counts, int accumulate) {
    __shared__ CauseLds L;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x)
        cause_row(cause + (uint64_t)(uint32_t)e * ld, n, e, pred_off, pred_key, pred_pos, pred_nkey, counts,
                  accumulate, L);
}

// ---- cause counts from winner positions --------------------------------------
// Rows of one byte per sample: the position of the entry that won the event's
// max (0xFF: none).  Counted by position in LDS (slot 0: none, 1 + k: entry
// k), then each entry's count moves to the first entry of the same source, so
// the counts equal cause_kernel's on the cause events.
constexpr int kPosRegs = 8;  // events with at most this many entries count in registers

template <int NP>
__device__ __forceinline__ void count_pos_regs(const uint8_t *__restrict__ row, int32_t n, uint32_t *c_lds) {
    uint32_t cnt[NP + 1];
#pragma unroll
    for (int k = 0; k <= NP; ++k) cnt[k] = 0;
    auto count = [&](uint32_t b) {
        cnt[0] += b == 0xFFu ? 1u : 0u;
#pragma unroll
        for (int k = 0; k < NP; ++k) cnt[1 + k] += b == (uint32_t)k ? 1u : 0u;
    };
    auto count4 = [&](uint32_t w) {
#pragma unroll
        for (int j = 0; j < 4; ++j) count((w >> (8 * j)) & 0xFFu);
    };
    // 16-byte loads over the row's aligned middle, single bytes around it
    const int32_t head = min(n, (int32_t)((16u - ((uint32_t)(uintptr_t)row & 15u)) & 15u));
    const int32_t n_vec = (n - head) / 16;
    const uint4 *vrow = reinterpret_cast<const uint4 *>(row + head);
    for (int32_t s = threadIdx.x; s < head; s += kBlock) count(row[s]);
    for (int32_t v = threadIdx.x; v < n_vec; v += kBlock) {
        const uint4 w = vrow[v];
        count4(w.x);
        count4(w.y);
        count4(w.z);
        count4(w.w);
    }
    for (int32_t s = head + 16 * n_vec + (int32_t)threadIdx.x; s < n; s += kBlock) count(row[s]);
#pragma unroll
    for (int k = 0; k <= NP; ++k) {
        uint32_t t = cnt[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(&c_lds[k], t);
    }
}

// The counts of event e from its row of n positions, by a block of kBlock
// threads; c_lds holds at least 256 counters.
__device__ __forceinline__ void cause_pos_row(const uint8_t *__restrict__ row, int32_t n, int32_t e,
                                              const int32_t *__restrict__ pred_off,
                                              const int32_t *__restrict__ pred_first, int64_t *__restrict__ counts,
                                              int accumulate, uint32_t *c_lds) {
    {
        const int32_t p0 = pred_off[e], np = pred_off[e + 1] - p0;  // np <= 254
        int64_t *out = counts + (int64_t)p0 + e;                   // [1 + np]
        for (int i = threadIdx.x; i <= np; i += kBlock) c_lds[i] = 0;
        __syncthreads();
        switch (np) {
            case 0: count_pos_regs<0>(row, n, c_lds); break;
            case 1: count_pos_regs<1>(row, n, c_lds); break;
            case 2: count_pos_regs<2>(row, n, c_lds); break;
            case 3: count_pos_regs<3>(row, n, c_lds); break;
            case 4: count_pos_regs<4>(row, n, c_lds); break;
            case 5: count_pos_regs<5>(row, n, c_lds); break;
            case 6: count_pos_regs<6>(row, n, c_lds); break;
            case 7: count_pos_regs<7>(row, n, c_lds); break;
            case 8: count_pos_regs<8>(row, n, c_lds); break;
            default: {
                // a wave's positions take few distinct values: one atomic per value
                const int32_t n_up = (n + kBlock - 1) / kBlock * kBlock;  // every lane takes part in the ballots
                for (int32_t s = threadIdx.x; s < n_up; s += kBlock) {
                    const uint32_t b = s < n ? row[s] : 0xFFFFu;  // past the row: counted nowhere
                    const int32_t slot = b == 0xFFu ? 0 : (b < (uint32_t)np ? 1 + (int32_t)b : -1);
                    bool pending = slot >= 0;
                    while (__ballot(pending)) {
                        const uint64_t mm = __ballot(pending);
                        const int leader = __ffsll((long long)mm) - 1;
                        const int32_t s0 = __shfl(slot, leader);
                        const uint64_t same = __ballot(pending && slot == s0);
                        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&c_lds[s0], (uint32_t)__popcll(same));
                        pending = pending && slot != s0;
                    }
                }
            }
        }
        __syncthreads();
        // an entry whose source occurs earlier in the list counts for that first entry
        for (int k = threadIdx.x; k < np; k += kBlock) {
            const int32_t f = pred_first[p0 + k];
            if (f != k) {
                const uint32_t v = c_lds[1 + k];
                c_lds[1 + k] = 0;
                if (v) atomicAdd(&c_lds[1 + f], v);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i <= np; i += kBlock) out[i] = (accumulate ? out[i] : 0) + (int64_t)c_lds[i];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) cause_pos_kernel(const uint8_t *__restrict__ pos, int32_t n, uint32_t ld,
                                                           int32_t E, const int32_t *__restrict__ pred_off,
                                                           const int32_t *__restrict__ pred_first,
                                                           int64_t *__restrict__ counts, int accumulate) {
    __shared__ uint32_t c_lds[1 + 256];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x)
        cause_pos_row(pos + (uint64_t)(uint32_t)e * ld, n, e, pred_off, pred_first, counts, accumulate, c_lds);
}

// ---- several DAGs' statistics in one launch ----------------------------------
// The set holding global row g: the last with row0 <= g (sets in row order).
__device__ __forceinline__ int32_t find_set(const RowSet *__restrict__ sets, int32_t n_sets, int32_t g) {
    int32_t lo = 0, hi = n_sets - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (sets[mid].row0 <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kQBlock) quantile_sets_kernel(const RowSet *__restrict__ sets, int32_t n_sets,
                                                                int32_t rows) {
    __shared__ QuantileLds L;
    for (int32_t g = blockIdx.x; g < rows; g += gridDim.x) {
        const RowSet &S = sets[find_set(sets, n_sets, g)];
        const int32_t e = g - S.row0;
        quantile_row<true>(S.realized + (uint64_t)(uint32_t)e * S.ld, S.n, S.qa,
                           S.quant + (uint64_t)(uint32_t)e * (uint32_t)S.qa.n_q, S.mean + e, L);
    }
}

__global__ void __launch_bounds__(kBlock) mean_sets_kernel(const RowSet *__restrict__ sets, int32_t n_sets,
                                                           int32_t rows) {
    __shared__ DD part[kWaves];
    __shared__ double naive[kWaves];
    for (int32_t g = blockIdx.x; g < rows; g += gridDim.x) {
        const RowSet &S = sets[find_set(sets, n_sets, g)];
        const int32_t e = g - S.row0;
        mean_row(S.realized + (uint64_t)(uint32_t)e * S.ld, S.n, S.mean + e, part, naive);
    }
}

__global__ void __launch_bounds__(kBlock) cause_sets_kernel(const RowSet *__restrict__ sets, int32_t n_sets,
                                                            int32_t rows) {
    __shared__ CauseLds L;
    for (int32_t g = blockIdx.x; g < rows; g += gridDim.x) {
        const RowSet &S = sets[find_set(sets, n_sets, g)];
        if (!S.counts) continue;  // uniform over the block
        const int32_t e = g - S.row0;
        if (S.cause_pos)
            cause_pos_row(S.cause_pos + (uint64_t)(uint32_t)e * S.ld, S.n, e, S.pred_off, S.pred_first, S.counts, 0,
                          L.c);
        else
            cause_row(S.cause + (uint64_t)(uint32_t)e * S.ld, S.n, e, S.pred_off, S.pred_key, S.pred_pos, S.pred_nkey,
                      S.counts, 0, L);
    }
}

// ---- statistics over sample chunks ------------------------------------------
// The histogram bin of v in its event's range [lo, lo + max_delay]
// (inv_width = n_bins / max_delay): monotone in v.
__device__ __forceinline__ int32_t value_bin(double v, double lo, double inv_width, int32_t n_bins) {
    double f = (v - lo) * inv_width;
    f = f < 0.0 ? 0.0 : f;
    return f < (double)n_bins ? (int32_t)f : n_bins - 1;
}

// An event's running statistics over chunks of samples: the exactly rounded
// sum (double-double, with the plain sum for non-finite rows) and the
// extremes, merged in chunk order; and a histogram of `n_bins` bins over the
// event's range [earliest, earliest + max_delay], where every realized time
// lies.  Quantiles interpolate inside a bin, so each is within one bin width
// of numpy's value; the mean and the extremes are exact.
__global__ void __launch_bounds__(kBlock) chunk_accumulate_kernel(const double *__restrict__ realized, int32_t m,
                                                                  uint32_t ld, int32_t E,
                                                                  const double *__restrict__ earliest,
                                                                  double inv_width, int32_t n_bins, int first,
                                                                  ChunkStat *__restrict__ st,
                                                                  uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[kMaxChunkBins];
    __shared__ Extremes ext[kWaves];
    __shared__ DD part[kWaves];
    __shared__ double naive[kWaves];
    const int lane = threadIdx.x & 63;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const double *row = realized + (uint64_t)(uint32_t)e * ld;
        const double lo = earliest[e];
        for (int i = threadIdx.x; i < n_bins; i += kBlock) h[i] = 0;
        __syncthreads();
        Extremes x{~0ull, 0ull, 0u, 0u, 0u};
        DD acc{0.0, 0.0};
        double plain = 0.0;
        // every lane takes part in each step (the histogram adds are wave-wide);
        // four loads per lane in flight while a whole block of four fits
        auto visit = [&](double v, bool in) {
            const bool ok = in && v == v;
            if (in) {
                acc = dd_add(acc, v);
                plain += v;
                if (!ok) ++x.nans;
            }
            if (ok) {
                const uint64_t key = order_key(v);
                x = merge_extremes(x, Extremes{key, key, 1u, 1u, 0u});
            }
            wave_hist_add(h, ok, (uint32_t)value_bin(v, lo, inv_width, n_bins), lane);
        };
        int32_t s0 = 0;
        for (; s0 + 4 * kBlock <= m; s0 += 4 * kBlock) {
            const double *p = row + s0 + threadIdx.x;
            const double v0 = p[0], v1 = p[kBlock], v2 = p[2 * kBlock], v3 = p[3 * kBlock];
            visit(v0, true);
            visit(v1, true);
            visit(v2, true);
            visit(v3, true);
        }
        for (; s0 < m; s0 += kBlock) {
            const int32_t s = s0 + (int32_t)threadIdx.x;
            const bool in = s < m;
            visit(in ? row[s] : 0.0, in);
        }
        x = wave_merge_extremes(x);
        wave_sum(acc, plain);
        if (lane == 0) {
            ext[threadIdx.x >> 6] = x;
            part[threadIdx.x >> 6] = acc;
            naive[threadIdx.x >> 6] = plain;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            DD sum = dd_add(dd_add(part[0], part[1]), dd_add(part[2], part[3]));
            double nv = (naive[0] + naive[1]) + (naive[2] + naive[3]);
            Extremes x0 = merge_extremes(merge_extremes(ext[0], ext[1]), merge_extremes(ext[2], ext[3]));
            if (!first) {
                const ChunkStat o = st[e];
                sum = dd_add(DD{o.sum_hi, o.sum_lo}, sum);
                nv = o.naive + nv;
                x0 = merge_extremes(Extremes{o.mn, o.mx, o.cmn, o.cmx, o.nans}, x0);
            }
            ChunkStat c;
            c.sum_hi = sum.hi;
            c.sum_lo = sum.lo;
            c.naive = nv;
            c.mn = x0.mn;
            c.mx = x0.mx;
            c.cmn = x0.cmn;
            c.cmx = x0.cmx;
            c.nans = x0.nans;
            c.pad = 0;
            st[e] = c;
        }
        // the first chunk writes every bin; later ones touch only the bins they
        // hit (a row saturated at its cap hits one), not 16 KiB per row
        uint32_t *gh = hist + (uint64_t)(uint32_t)e * (uint32_t)n_bins;
        for (int i = threadIdx.x; i < n_bins; i += kBlock) {
            const uint32_t c = h[i];
            if (first)
                gh[i] = c;
            else if (c)
                gh[i] += c;
        }
        __syncthreads();
    }
}

// Means and quantiles from the chunk statistics of n samples (numpy's
// "linear" method on order statistics estimated inside their histogram bin).
__global__ void __launch_bounds__(kBlock) chunk_finish_kernel(int64_t n, int32_t E, QuantileArgs qa,
                                                              const double *__restrict__ earliest, double width,
                                                              int32_t n_bins, const ChunkStat *__restrict__ st,
                                                              const uint32_t *__restrict__ hist,
                                                              double *__restrict__ mean, double *__restrict__ out) {
    __shared__ uint64_t seg[kBlock];  // bins per thread: n_bins / kBlock (a power of two)
    __shared__ double stat[kMaxRanks];
    const int per = n_bins / kBlock;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const ChunkStat c = st[e];
        const uint32_t *h = hist + (uint64_t)(uint32_t)e * (uint32_t)n_bins;
        uint64_t t = 0;
        for (int i = 0; i < per; ++i) t += h[threadIdx.x * per + i];
        seg[threadIdx.x] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            if (mean) mean[e] = (__builtin_isfinite(c.naive) ? c.sum_hi : c.naive) / (double)n;
            const double lo = earliest[e];
            const int64_t valid = n - (int64_t)c.nans;
            for (int r = 0; r < qa.n_ranks; ++r) {
                const int64_t k = qa.rank[r];
                double v;
                if (c.nans) {
                    v = __builtin_nan("");
                } else if (k >= valid - (int64_t)c.cmx) {
                    v = key_value(c.mx);
                } else if (k < (int64_t)c.cmn) {
                    v = key_value(c.mn);
                } else {
                    int64_t cum = 0;
                    int g = 0;
                    while (g < kBlock - 1 && cum + (int64_t)seg[g] <= k) cum += (int64_t)seg[g++];
                    int b = g * per;
                    while (b < (g + 1) * per - 1 && cum + (int64_t)h[b] <= k) cum += (int64_t)h[b++];
                    const double inside = ((double)(k - cum) + 0.5) / (double)(h[b] ? h[b] : 1u);
                    v = lo + ((double)b + inside) * width;
                    const double vmin = key_value(c.mn), vmax = key_value(c.mx);
                    v = v < vmin ? vmin : (v > vmax ? vmax : v);
                }
                stat[r] = v;
            }
            if (out)
                for (int j = 0; j < qa.n_q; ++j)
                    out[(uint64_t)(uint32_t)e * (uint32_t)qa.n_q + j] =
                        c.nans ? __builtin_nan("") : lerp_numpy(stat[qa.lo[j]], stat[qa.hi[j]], qa.gamma[j]);
        }
        __syncthreads();
    }
}

// ---- exact order statistics over further passes of the chunks ---------------
// The sub-bin of a value in a refined region: even splits of its value range,
// or of its key range once it spans few representable doubles.  Monotone in v.
__device__ __forceinline__ int32_t exact_sub_bin(const ExactTarget &t, double v, uint64_t key) {
    if (t.inv != 0.0) {
        double f = (v - t.a) * t.inv;
        f = f < 0.0 ? 0.0 : f;
        return f < (double)kExactSubBins ? (int32_t)f : kExactSubBins - 1;
    }
    return (int32_t)((key - t.klo) >> t.shift);
}

// Targets from the chunks' histograms: as chunk_finish_kernel places a rank,
// but keeping its bin and rank inside the bin instead of interpolating.
__global__ void __launch_bounds__(kBlock) exact_init_kernel(int64_t n, int32_t E, QuantileArgs qa,
                                                            const ChunkStat *__restrict__ st,
                                                            const uint32_t *__restrict__ hist, int32_t n_bins,
                                                            const double *__restrict__ earliest, double width,
                                                            ExactWork w) {
    __shared__ uint64_t seg[kBlock];
    const int per = n_bins / kBlock;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const ChunkStat c = st[e];
        const uint32_t *h = hist + (uint64_t)(uint32_t)e * (uint32_t)n_bins;
        uint64_t sum = 0;
        for (int i = 0; i < per; ++i) sum += h[threadIdx.x * per + i];
        seg[threadIdx.x] = sum;
        __syncthreads();
        if (threadIdx.x == 0) {
            const double lo = earliest[e];
            const int64_t valid = n - (int64_t)c.nans;
            for (int r = 0; r < w.n_ranks; ++r) {
                ExactTarget t{};
                t.slot = -1;
                t.mode = kExactResolved;
                const int64_t k = r < qa.n_ranks ? qa.rank[r] : 0;
                if (c.nans) {
                    t.value = __builtin_nan("");
                } else if (k >= valid - (int64_t)c.cmx) {
                    t.value = key_value(c.mx);
                } else if (k < (int64_t)c.cmn) {
                    t.value = key_value(c.mn);
                } else {
                    int64_t cum = 0;
                    int g = 0;
                    while (g < kBlock - 1 && cum + (int64_t)seg[g] <= k) cum += (int64_t)seg[g++];
                    int b = g * per;
                    while (b < (g + 1) * per - 1 && cum + (int64_t)h[b] <= k) cum += (int64_t)h[b++];
                    t.bin = b;
                    t.rank = k - cum;
                    t.count = h[b];
                    t.klo = 0;
                    t.khi = ~0ull;
                    t.a = lo + (double)b * width;
                    t.inv = (double)kExactSubBins / width;
                    t.mode = kExactWait;
                    if (t.rank < 0 || t.rank >= (int64_t)t.count) {
                        w.counters[3] = 1;  // the histogram does not hold the rank
                        t.mode = kExactResolved;
                        t.value = __builtin_nan("");
                    }
                }
                w.targets[(uint64_t)(uint32_t)e * (uint32_t)w.n_ranks + r] = t;
            }
        }
        __syncthreads();
    }
}

// Hands out the next pass's work: a waiting region with at most c_max values
// gets candidate space (while the buffer lasts), otherwise a refine slot
// (while slots last), otherwise waits for the pass after.  Counts the
// unresolved targets.
__global__ void __launch_bounds__(kBlock) exact_assign_kernel(int32_t E, ExactWork w) {
    __shared__ int32_t new_slot[kMaxRanks];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        ExactTarget *tg = w.targets + (uint64_t)(uint32_t)e * (uint32_t)w.n_ranks;
        if (threadIdx.x == 0) {
            for (int r = 0; r < w.n_ranks; ++r) {
                new_slot[r] = -1;
                ExactTarget t = tg[r];
                if (t.mode == kExactResolved) continue;
                t.mode = kExactWait;
                t.slot = -1;
                if (t.count <= (uint32_t)w.c_max) {
                    const uint64_t off = atomicAdd((unsigned long long *)&w.counters[0], (unsigned long long)t.count);
                    if (off + t.count <= (uint64_t)w.cand_cap) {
                        t.mode = kExactCollect;
                        t.off = (uint32_t)off;
                        t.fill = 0;
                    }
                }
                if (t.mode == kExactWait) {
                    const uint64_t sl = atomicAdd((unsigned long long *)&w.counters[1], 1ull);
                    if (sl < (uint64_t)w.n_slots) {
                        t.mode = kExactRefine;
                        t.slot = (int32_t)sl;
                        new_slot[r] = (int32_t)sl;
                    }
                }
                atomicAdd((unsigned long long *)&w.counters[2], 1ull);
                tg[r] = t;
            }
        }
        __syncthreads();
        for (int r = 0; r < w.n_ranks; ++r) {
            const int32_t sl = new_slot[r];
            if (sl < 0) continue;
            const uint64_t base = (uint64_t)(uint32_t)sl * kExactSubBins;
            for (int i = threadIdx.x; i < kExactSubBins; i += kBlock) {
                w.slot_cnt[base + i] = 0u;
                w.slot_min[base + i] = ~0ull;
                w.slot_max[base + i] = 0ull;
            }
        }
        __syncthreads();
    }
}

__global__ void exact_reset_kernel(uint64_t *counters, int all) {
    counters[0] = 0;
    counters[1] = 0;
    counters[2] = 0;
    if (all) counters[3] = 0;
}

// One chunk of a pass: every value of an open region is gathered (a wave's
// matching lanes take consecutive places) or counted in its sub-bin.
__global__ void __launch_bounds__(kBlock) exact_pass_kernel(const double *__restrict__ realized, int32_t m,
                                                            uint32_t ld, int32_t E,
                                                            const double *__restrict__ earliest, double inv_width,
                                                            int32_t n_bins, ExactWork w) {
    __shared__ ExactTarget tg[kMaxRanks];
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        ExactTarget *gt = w.targets + (uint64_t)(uint32_t)e * (uint32_t)w.n_ranks;
        if ((int)threadIdx.x < w.n_ranks) tg[threadIdx.x] = gt[threadIdx.x];
        __syncthreads();
        uint32_t open = 0;
        for (int r = 0; r < w.n_ranks; ++r)
            if (tg[r].mode == kExactCollect || tg[r].mode == kExactRefine) open |= 1u << r;
        if (open) {
            const double *row = realized + (uint64_t)(uint32_t)e * ld;
            const double lo = earliest[e];
            for (int32_t s0 = 0; s0 < m; s0 += kBlock) {  // every lane takes part in each step (ballots)
                const int32_t s = s0 + (int32_t)threadIdx.x;
                const bool in = s < m;
                const double v = in ? row[s] : 0.0;
                const bool ok = in && v == v;
                const int32_t vb = value_bin(v, lo, inv_width, n_bins);
                const uint64_t key = order_key(v);
                for (uint32_t o = open; o; o &= o - 1) {
                    const int r = __ffs((int)o) - 1;
                    const ExactTarget &t = tg[r];
                    const bool mem = ok && vb == t.bin && key >= t.klo && key <= t.khi;
                    if (t.mode == kExactCollect) {
                        const uint64_t mm = __ballot(mem);
                        if (mm == 0) continue;
                        const int leader = __ffsll((long long)mm) - 1;
                        uint32_t base = 0;
                        if (lane == leader) base = atomicAdd(&gt[r].fill, (uint32_t)__popcll(mm));
                        base = __shfl(base, leader);
                        const uint32_t pos = base + (uint32_t)__popcll(mm & below);
                        if (mem) {
                            if (pos < t.count)
                                w.cand[(uint64_t)t.off + pos] = v;
                            else
                                w.counters[3] = 1;  // more values than the histogram counted
                        }
                    } else if (mem) {
                        const uint64_t i = (uint64_t)(uint32_t)t.slot * kExactSubBins + (uint32_t)exact_sub_bin(t, v, key);
                        atomicAdd(&w.slot_cnt[i], 1u);
                        atomicMin((unsigned long long *)&w.slot_min[i], (unsigned long long)key);
                        atomicMax((unsigned long long *)&w.slot_max[i], (unsigned long long)key);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// After a pass: a gathered region is sorted (bitonic, in LDS) and its rank
// read off; a refined region becomes the sub-bin holding its rank, bounded by
// that sub-bin's smallest and largest key.
__global__ void __launch_bounds__(kBlock) exact_resolve_kernel(int32_t E, ExactWork w) {
    __shared__ uint64_t keys[kExactMaxCollect];
    for (int32_t e = blockIdx.x; e < E; e += gridDim.x) {
        ExactTarget *gt = w.targets + (uint64_t)(uint32_t)e * (uint32_t)w.n_ranks;
        for (int r = 0; r < w.n_ranks; ++r) {
            const ExactTarget t = gt[r];
            if (t.mode == kExactCollect) {
                const uint32_t c = t.count;
                uint32_t p2 = 1;
                while (p2 < c) p2 <<= 1;
                for (uint32_t i = threadIdx.x; i < p2; i += kBlock)
                    keys[i] = i < c ? order_key(w.cand[(uint64_t)t.off + i]) : ~0ull;
                __syncthreads();
                for (uint32_t k = 2; k <= p2; k <<= 1)
                    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                        for (uint32_t i = threadIdx.x; i < p2; i += kBlock) {
                            const uint32_t x = i ^ j;
                            if (x > i) {
                                const uint64_t a = keys[i], b = keys[x];
                                if ((a > b) == ((i & k) == 0)) {
                                    keys[i] = b;
                                    keys[x] = a;
                                }
                            }
                        }
                        __syncthreads();
                    }
                if (threadIdx.x == 0) {
                    ExactTarget u = t;
                    if (t.fill != c) {
                        w.counters[3] = 1;
                        u.value = __builtin_nan("");
                    } else {
                        u.value = key_value(keys[t.rank]);
                    }
                    u.mode = kExactResolved;
                    gt[r] = u;
                }
                __syncthreads();
            } else if (t.mode == kExactRefine && threadIdx.x == 0) {
                const uint64_t base = (uint64_t)(uint32_t)t.slot * kExactSubBins;
                int64_t cum = 0;
                int j = 0;
                while (j < kExactSubBins - 1 && cum + (int64_t)w.slot_cnt[base + j] <= t.rank) cum += w.slot_cnt[base + j++];
                ExactTarget u = t;
                u.rank = t.rank - cum;
                u.count = w.slot_cnt[base + j];
                u.klo = w.slot_min[base + j];
                u.khi = w.slot_max[base + j];
                u.slot = -1;
                if (u.rank < 0 || u.rank >= (int64_t)u.count || u.klo > u.khi) {
                    w.counters[3] = 1;  // the sub-bins do not hold the rank
                    u.mode = kExactResolved;
                    u.value = __builtin_nan("");
                } else if (u.klo == u.khi) {
                    u.mode = kExactResolved;
                    u.value = key_value(u.klo);
                } else {
                    u.mode = kExactWait;
                    const uint64_t span = u.khi - u.klo;
                    const double va = key_value(u.klo), vb = key_value(u.khi);
                    if (span > (1ull << 20) && vb - va > 0.0) {
                        u.a = va;
                        u.inv = (double)kExactSubBins / (vb - va);
                        u.shift = 0;
                    } else {
                        u.inv = 0.0;
                        u.shift = 0;
                        while ((span >> u.shift) >= (uint64_t)kExactSubBins) ++u.shift;
                    }
                }
                gt[r] = u;
            }
        }
        __syncthreads();
    }
}

__global__ void exact_finish_kernel(int32_t E, QuantileArgs qa, const ChunkStat *__restrict__ st, ExactWork w,
                                    double *__restrict__ out) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const ExactTarget *t = w.targets + (uint64_t)(uint32_t)e * (uint32_t)w.n_ranks;
    const bool nan_row = st[e].nans != 0;
    for (int j = 0; j < qa.n_q; ++j)
        out[(uint64_t)(uint32_t)e * (uint32_t)qa.n_q + j] =
            nan_row ? __builtin_nan("") : lerp_numpy(t[qa.lo[j]].value, t[qa.hi[j]].value, qa.gamma[j]);
}

// ---- rows kept whole across the chunks of a request -------------------------
__global__ void counters_clear_kernel(uint64_t *c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

// One event per thread: is any rank of the n samples outside the masses of
// the extremes, as far as the first chunk's m samples tell?  A mass counts
// only where its extreme is the event's bound; its fraction f is widened by
// 4 sqrt(max(f (1 - f), 1/m) / m) + 2/m.  Listed in the order the atomic
// hands out places (the order does not change any result).
__global__ void __launch_bounds__(256) watch_select_kernel(int64_t n, int32_t m, int32_t E, QuantileArgs qa,
                                                           const ChunkStat *__restrict__ st,
                                                           const double *__restrict__ ranges, WatchWork w) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const ChunkStat c = st[e];
    if (c.nans || c.cmn == 0) return;  // a row with NaN has NaN quantiles, resolved at once
    const double inv_m = 1.0 / (double)m;
    const double f_lo = key_value(c.mn) == ranges[2 * (int64_t)e] ? (double)c.cmn * inv_m : 0.0;
    const double f_hi = key_value(c.mx) == ranges[2 * (int64_t)e + 1] ? (double)c.cmx * inv_m : 0.0;
    auto widen = [&](double f) {
        const double v = f * (1.0 - f);
        return 4.0 * sqrt((v > inv_m ? v : inv_m) * inv_m) + 2.0 * inv_m;
    };
    const double k_lo = (f_lo - widen(f_lo)) * (double)n;
    const double k_hi = (1.0 - f_hi + widen(f_hi)) * (double)n;
    // ranks 0 and n - 1 always lie on the masses (each holds a sample at least)
    bool open = false;
    for (int r = 0; r < qa.n_ranks; ++r) {
        const double k = (double)qa.rank[r];
        open = open || (k > 0.0 && k < (double)(n - 1) && k >= k_lo && k < k_hi);
    }
    if (!open) return;
    const uint64_t i = atomicAdd((unsigned long long *)w.count, 1ull);
    if (i < (uint64_t)w.cap) w.list[i] = e;
}

__device__ __forceinline__ uint32_t watch_rows(const WatchWork &w) {
    const uint64_t c = *w.count;
    return c < (uint64_t)w.cap ? (uint32_t)c : (uint32_t)w.cap;
}

__global__ void __launch_bounds__(256) watch_copy_kernel(const double *__restrict__ realized, int32_t m, uint32_t ld,
                                                         int64_t c0, WatchWork w) {
    const uint32_t rows = watch_rows(w);
    for (uint32_t i = blockIdx.x; i < rows; i += gridDim.x) {
        const double *src = realized + (uint64_t)(uint32_t)w.list[i] * ld;
        double *dst = w.store + (uint64_t)i * (uint64_t)w.ld + (uint64_t)c0;
        for (int32_t s = threadIdx.x; s < m; s += blockDim.x) dst[s] = src[s];
    }
}

// The targets of kept rows need no selection: resolved (their quantiles are
// written over exact_finish_kernel's by launch_watch_scatter).
__global__ void __launch_bounds__(256) watch_mark_kernel(WatchWork w, ExactWork x) {
    const uint32_t rows = watch_rows(w);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) {
        ExactTarget *t = x.targets + (uint64_t)(uint32_t)w.list[i] * (uint32_t)x.n_ranks;
        for (int r = 0; r < x.n_ranks; ++r) {
            t[r].mode = kExactResolved;
            t[r].slot = -1;
            t[r].value = 0.0;
        }
    }
}

__global__ void __launch_bounds__(256) watch_scatter_kernel(WatchWork w, int32_t rows, int32_t n_q,
                                                            double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * n_q) return;
    const int32_t r = (int32_t)(i / n_q), j = (int32_t)(i % n_q);
    out[(uint64_t)(uint32_t)w.list[r] * (uint32_t)n_q + j] = w.wq[i];
}

}  // namespace

hipError_t launch_counters_clear(uint64_t *c, int n, hipStream_t stream) {
    hipLaunchKernelGGL(counters_clear_kernel, dim3(1), dim3(64), 0, stream, c, n);
    return hipGetLastError();
}

hipError_t launch_watch_select(int64_t n, int32_t m, int32_t E, const QuantileArgs &qa, const ChunkStat *st,
                               const double *ranges, const WatchWork &w, hipStream_t stream) {
    if (E > 0 && m > 0 && w.cap > 0)
        hipLaunchKernelGGL(watch_select_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, stream, n, m, E, qa,
                           st, ranges, w);
    return hipGetLastError();
}

hipError_t launch_watch_copy(const double *realized, int32_t m, int32_t ld, int64_t c0, const WatchWork &w,
                             hipStream_t stream) {
    if (m > 0 && w.cap > 0)
        hipLaunchKernelGGL(watch_copy_kernel, dim3((unsigned)(w.cap < 1024 ? w.cap : 1024)), dim3(256), 0, stream, realized, m,
                           (uint32_t)ld, c0, w);
    return hipGetLastError();
}

hipError_t launch_watch_scatter(const WatchWork &w, int32_t rows, int32_t n_q, double *out, hipStream_t stream) {
    const int64_t total = (int64_t)rows * n_q;
    if (total > 0)
        hipLaunchKernelGGL(watch_scatter_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, w, rows,
                           n_q, out);
    return hipGetLastError();
}

hipError_t launch_chunk_accumulate(const double *realized, int32_t m, int32_t ld, int32_t E, const double *earliest,
                                   double max_delay, int32_t n_bins, int first, ChunkStat *st, uint32_t *hist,
                                   hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(chunk_accumulate_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, realized, m,
                           (uint32_t)ld, E, earliest, (double)n_bins / max_delay, n_bins, first, st, hist);
    return hipGetLastError();
}

hipError_t launch_chunk_finish(int64_t n, int32_t E, const QuantileArgs &qa, const double *earliest, double max_delay,
                               int32_t n_bins, const ChunkStat *st, const uint32_t *hist, double *mean, double *out,
                               hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(chunk_finish_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, n, E, qa, earliest,
                           max_delay / (double)n_bins, n_bins, st, hist, mean, out);
    return hipGetLastError();
}

hipError_t launch_exact_init(int64_t n, int32_t E, const QuantileArgs &qa, const ChunkStat *st, const uint32_t *hist,
                             int32_t n_bins, const double *earliest, double max_delay, const ExactWork &w,
                             const WatchWork *watch, hipStream_t stream) {
    hipLaunchKernelGGL(exact_reset_kernel, dim3(1), dim3(1), 0, stream, w.counters, 1);
    if (E > 0) {
        hipLaunchKernelGGL(exact_init_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, n, E, qa, st, hist, n_bins,
                           earliest, max_delay / (double)n_bins, w);
        if (watch && watch->cap > 0)
            hipLaunchKernelGGL(watch_mark_kernel, dim3((unsigned)((watch->cap + 255) / 256 < 1024 ? (watch->cap + 255) / 256 : 1024)), dim3(256), 0,
                               stream, *watch, w);
        hipLaunchKernelGGL(exact_assign_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, E, w);
    }
    return hipGetLastError();
}

hipError_t launch_exact_pass(const double *realized, int32_t m, int32_t ld, int32_t E, const double *earliest,
                             double max_delay, int32_t n_bins, const ExactWork &w, hipStream_t stream) {
    if (E > 0 && m > 0)
        hipLaunchKernelGGL(exact_pass_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, realized, m, (uint32_t)ld, E,
                           earliest, (double)n_bins / max_delay, n_bins, w);
    return hipGetLastError();
}

hipError_t launch_exact_round_end(int32_t E, const ExactWork &w, hipStream_t stream) {
    if (E > 0) hipLaunchKernelGGL(exact_resolve_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, E, w);
    hipLaunchKernelGGL(exact_reset_kernel, dim3(1), dim3(1), 0, stream, w.counters, 0);
    if (E > 0) hipLaunchKernelGGL(exact_assign_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, E, w);
    return hipGetLastError();
}

hipError_t launch_exact_finish(int32_t E, const QuantileArgs &qa, const ChunkStat *st, const ExactWork &w,
                               double *out, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(exact_finish_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, stream, E, qa, st, w,
                           out);
    return hipGetLastError();
}

hipError_t launch_set_stats(const RowSet *sets, int32_t n_sets, int32_t rows, bool quantiles, bool counts,
                            hipStream_t stream) {
    if (rows <= 0 || n_sets <= 0) return hipSuccess;
    if (quantiles)
        hipLaunchKernelGGL(quantile_sets_kernel, dim3((unsigned)rows), dim3(kQBlock), 0, stream, sets, n_sets, rows);
    else
        hipLaunchKernelGGL(mean_sets_kernel, dim3((unsigned)rows), dim3(kBlock), 0, stream, sets, n_sets, rows);
    if (counts) hipLaunchKernelGGL(cause_sets_kernel, dim3((unsigned)rows), dim3(kBlock), 0, stream, sets, n_sets, rows);
    return hipGetLastError();
}

hipError_t launch_event_mean(const double *realized, int32_t n, int32_t ld, int32_t E, double *mean,
                             hipStream_t stream) {
    if (E > 0) hipLaunchKernelGGL(mean_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, realized, n, (uint32_t)ld, E, mean);
    return hipGetLastError();
}

hipError_t launch_event_quantiles(const double *realized, int32_t n, int32_t ld, int32_t E, const QuantileArgs &qa,
                                  double *out, double *mean, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(mean ? quantile_kernel<true> : quantile_kernel<false>, dim3((unsigned)E), dim3(kQBlock), 0,
                           stream, realized, n, (uint32_t)ld, E, qa, out, mean);
    return hipGetLastError();
}

hipError_t launch_cause_counts(const int32_t *cause, int32_t n, int32_t ld, int32_t E, const int32_t *pred_off,
                               const int32_t *pred_key, const int32_t *pred_pos, const int32_t *pred_nkey,
                               int64_t *counts, int accumulate, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(cause_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, cause, n, (uint32_t)ld, E, pred_off,
                           pred_key, pred_pos, pred_nkey, counts, accumulate);
    return hipGetLastError();
}

hipError_t launch_cause_pos_counts(const uint8_t *pos, int32_t n, int32_t ld, int32_t E, const int32_t *pred_off,
                                   const int32_t *pred_first, int64_t *counts, int accumulate, hipStream_t stream) {
    if (E > 0)
        hipLaunchKernelGGL(cause_pos_kernel, dim3((unsigned)E), dim3(kBlock), 0, stream, pos, n, (uint32_t)ld, E,
                           pred_off, pred_first, counts, accumulate);
    return hipGetLastError();
}

}  // namespace mcdp
