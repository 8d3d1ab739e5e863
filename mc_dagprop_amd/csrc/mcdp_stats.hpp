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
// One DAG's rows for statistics of several DAGs in one launch: rows
// realized[e * ld + s] and cause likewise (s < n, e < E), global rows
// row0 .. row0 + E - 1 (the sets in row order), outputs mean [E], quant
// [E][qa.n_q] and, when counts is not null, the cause counts (the layout of
// launch_cause_counts, from the DAG's predecessor table; from winner positions
// when cause_pos is not null).
struct RowSet {
    const double *realized;
    const int32_t *cause;
    const uint8_t *cause_pos;  // instead of cause when not null: winner positions (launch_cause_pos_counts)
    const int32_t *pred_first;
    double *mean;
    double *quant;
    int64_t *counts;
    const int32_t *pred_off, *pred_key, *pred_pos, *pred_nkey;
    int32_t n;
    uint32_t ld;
    int32_t E;
    int32_t row0;
    QuantileArgs qa;
};
// The statistics of launch_event_quantiles (quantiles: with the means) or
// launch_event_mean, and of launch_cause_counts (counts), for `rows` rows of
// n_sets sets in device memory; each row's results equal those kernels'.
hipError_t launch_set_stats(const RowSet *sets, int32_t n_sets, int32_t rows, bool quantiles, bool counts,
                            hipStream_t stream);
// The cause counts of launch_cause_counts from winner positions (one byte per
// sample: the winning entry's position in the event's list, 0xFF: none; events
// of at most 254 entries): pred_first[p0 + k] is the first entry of entry k's
// source in the event's list.
hipError_t launch_cause_pos_counts(const uint8_t *pos, int32_t n, int32_t ld, int32_t E, const int32_t *pred_off,
                                   const int32_t *pred_first, int64_t *counts, int accumulate, hipStream_t stream);
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
// mean[e] (exactly rounded sum / n) and, when out is not null, out[e * n_q + j]:
// numpy-linear quantiles from order statistics placed inside their bin (within
// one bin width).
hipError_t launch_chunk_finish(int64_t n, int32_t E, const QuantileArgs &qa, const double *earliest, double max_delay,
                               int32_t n_bins, const ChunkStat *st, const uint32_t *hist, double *mean, double *out,
                               hipStream_t stream);

// Exact order statistics of a request run in chunks, over further passes of
// the same chunks (their results are a function of the seeds).  Each (event,
// rank) target is a region of the event's values: those in histogram bin
// `bin` with klo <= order key <= khi, holding `count` values, the wanted one
// at `rank` among them.  A pass either gathers a region's values (at most
// c_max; then they are sorted and the rank read off) or counts them in 256
// sub-bins with each sub-bin's smallest and largest key (then the region
// shrinks to the sub-bin holding the rank).  Sub-bins split the region's value
// range evenly while it spans many representable doubles, then its key range
// (key offsets >> shift), so a region ends as one key or few enough values.
enum ExactMode : int32_t { kExactResolved = 0, kExactCollect = 1, kExactRefine = 2, kExactWait = 3 };
struct ExactTarget {
    uint64_t klo, khi;
    double a, inv;   // value sub-bins: (v - a) * inv, clamped to [0, 255]; inv == 0: (key - klo) >> shift
    int64_t rank;
    uint32_t count;
    int32_t bin;
    int32_t mode;    // ExactMode
    int32_t shift;
    int32_t slot;    // kExactRefine: the refine slot
    uint32_t off;    // kExactCollect: the first candidate
    uint32_t fill;   // kExactCollect: candidates gathered so far
    uint32_t pad;
    double value;    // kExactResolved: the order statistic
};
static_assert(sizeof(ExactTarget) == 80, "ExactTarget layout");
constexpr int kExactSubBins = 256;
constexpr int kExactMaxCollect = 4096;  // values gathered per region (sorted in LDS)
// Device state of the exact selection: targets [E * n_ranks], candidates, the
// refine slots' sub-bin counts, smallest and largest keys [n_slots * 256]
// each, and counters (candidates handed out, slots handed out, unresolved
// targets, an inconsistency flag).
struct ExactWork {
    ExactTarget *targets;
    double *cand;
    uint32_t *slot_cnt;
    uint64_t *slot_min, *slot_max;
    uint64_t *counters;  // [4]
    int32_t n_ranks;
    int32_t n_slots;
    int32_t c_max;
    uint32_t cand_cap;
};
// Rows kept whole for a request run in chunks.  After the first chunk, the
// events whose order statistics may leave the masses of their extremes are
// listed; each later chunk appends their samples to a row of n, so their
// quantiles come from the whole row (the one-piece kernel) and their targets
// need no further pass.  list [cap] event ids; *count the events chosen (past
// cap they are not kept and take further passes if they need them); store
// [cap][ld] with ld the request's samples; wq [cap][n_q] their quantiles.
struct WatchWork {
    int32_t *list;
    uint64_t *count;
    double *store;
    double *wq;
    int32_t cap;
    int64_t ld;
};
// c[0..n) = 0 (a kernel, so the clearing may sit in a captured graph).
hipError_t launch_counters_clear(uint64_t *c, int n, hipStream_t stream);
// Chooses the rows to keep from the first chunk's statistics st (m samples of
// n): an extreme's mass counts only where the extreme equals the event's bound
// (ranges[2e], ranges[2e + 1]: no later value can pass it), and each mass
// fraction is widened by four standard errors and two samples.
hipError_t launch_watch_select(int64_t n, int32_t m, int32_t E, const QuantileArgs &qa, const ChunkStat *st,
                               const double *ranges, const WatchWork &w, hipStream_t stream);
// Appends samples c0 .. c0 + m of the kept rows (realized[e * ld + s], s < m).
hipError_t launch_watch_copy(const double *realized, int32_t m, int32_t ld, int64_t c0, const WatchWork &w,
                             hipStream_t stream);
// out[e * n_q + j] = wq[i * n_q + j] for the first `rows` kept events e = list[i].
hipError_t launch_watch_scatter(const WatchWork &w, int32_t rows, int32_t n_q, double *out, hipStream_t stream);
// After the chunks' running statistics: targets from the histogram (ranks on
// an extreme's mass, or of a row with NaN, are resolved at once; so are those
// of kept rows, when watch is given), then modes handed out.  counters[2] =
// unresolved targets, counters[3] != 0 on an inconsistency.
hipError_t launch_exact_init(int64_t n, int32_t E, const QuantileArgs &qa, const ChunkStat *st, const uint32_t *hist,
                             int32_t n_bins, const double *earliest, double max_delay, const ExactWork &w,
                             const WatchWork *watch, hipStream_t stream);
// One chunk of a pass: rows realized[e][0..m) against every open target.
hipError_t launch_exact_pass(const double *realized, int32_t m, int32_t ld, int32_t E, const double *earliest,
                             double max_delay, int32_t n_bins, const ExactWork &w, hipStream_t stream);
// After a pass over every chunk: gathered regions sorted and read, refined
// regions shrunk; then modes handed out again.
hipError_t launch_exact_round_end(int32_t E, const ExactWork &w, hipStream_t stream);
// out[e * n_q + j] from the resolved order statistics (numpy's "linear").
hipError_t launch_exact_finish(int32_t E, const QuantileArgs &qa, const ChunkStat *st, const ExactWork &w,
                               double *out, hipStream_t stream);

}  // namespace mcdp
