// Launch interface of the per-event statistics kernels (mcdp_stats.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace mcdp {

constexpr int kMaxQuantiles = 8;
constexpr int kMaxRanks = 2 * kMaxQuantiles;
// Events with more predecessors than this count their causes with global
// atomics instead of LDS counters.
constexpr int kCauseLdsCounters = 4096;
// Events with more distinct predecessor sources than this search them in
// global memory instead of LDS.
constexpr int kCauseLdsKeys = 1024;

// Order statistics the quantiles need and how each quantile combines two of
// them (numpy's "linear" method; computed on the host in double).
struct QuantileArgs {
    int32_t n_q;
    int32_t n_ranks;
    int32_t rank[kMaxRanks];   // distinct 0-based ranks to select
    int32_t lo[kMaxQuantiles]; // index into rank[] of floor((n - 1) q), clamped
    int32_t hi[kMaxQuantiles]; // index into rank[] of the next order statistic
    double gamma[kMaxQuantiles];
};

hipError_t launch_event_mean(const double *realized, int32_t n, int32_t ld, int32_t E, double *mean,
                             hipStream_t stream);
// `mean` (nullable): the row means, computed in the quantile kernel's first pass.
hipError_t launch_event_quantiles(const double *realized, int32_t n, int32_t ld, int32_t E, const QuantileArgs &qa,
                                  double *out, double *mean, hipStream_t stream);
// accumulate != 0 adds to the counts already in `counts`.
hipError_t launch_cause_counts(const int32_t *cause, int32_t n, int32_t ld, int32_t E, const int32_t *pred_off,
                               const int32_t *pred_key, const int32_t *pred_pos, const int32_t *pred_nkey,
                               int64_t *counts, int accumulate, hipStream_t stream);

// Statistics of a request run in chunks of samples (serving).  Per event: the
// double-double sum, the plain sum, the extremes with their multiplicities and
// NaN count (ChunkStat), and a histogram of n_bins bins over
// [earliest, earliest + max_delay].
constexpr int kMaxChunkBins = 4096;  // a power of two and a multiple of the block size
struct ChunkStat {
    double sum_hi, sum_lo;
    double naive;
    uint64_t mn, mx;
    uint32_t cmn, cmx, nans, pad;
};
// first != 0 starts the sums; each call adds chunk rows realized[e][0..m).
hipError_t launch_chunk_accumulate(const double *realized, int32_t m, int32_t ld, int32_t E, const double *earliest,
                                   double max_delay, int32_t n_bins, int first, ChunkStat *st, uint32_t *hist,
                                   hipStream_t stream);
// mean[e] (exactly rounded sum / n) and out[e * n_q + j]: numpy-linear quantiles
// from order statistics placed inside their bin (within one bin width).
hipError_t launch_chunk_finish(int64_t n, int32_t E, const QuantileArgs &qa, const double *earliest, double max_delay,
                               int32_t n_bins, const ChunkStat *st, const uint32_t *hist, double *mean, double *out,
                               hipStream_t stream);

}  // namespace mcdp
