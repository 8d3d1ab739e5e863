// Per-event statistics of Monte Carlo results (C ABI in include/mcdp.h):
// the consumer side of run_many that the reference leaves to its callers
// (demo/monte_carlo.py:45-63 gathers every SimResult on the host and counts
// per event).  Here the statistics are reduced on the device, next to the
// event-major outputs, and only E-sized results leave it.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"
#include "mcdp_stats.hpp"

using namespace mcdp;

namespace {

// The predecessor table the cause counts need, on the device of `d`: per
// event, its distinct sources in ascending order with the first position of
// each in the event's predecessor list (a binary search maps a cause to its
// position), and per entry the first position of its source (maps a winner
// position to the same count).
int ensure_topo(const mcdp_propagator &P, DeviceCopy &d) {
    if (d.topo) return MCDP_OK;
    const int64_t np = (int64_t)P.preds.size();
    if (np > std::numeric_limits<int32_t>::max() - P.E - 1)
        return fail(MCDP_E_INVALID, "too many predecessor entries for cause counts");
    std::vector<int32_t> off((size_t)P.E + 1), key((size_t)np), pos((size_t)np), nkey((size_t)P.E), first((size_t)np);
    std::vector<std::pair<int32_t, int32_t>> tmp;
    for (int32_t e = 0; e < P.E; ++e) {
        const int64_t a = P.pred_off[(size_t)e], b = P.pred_off[(size_t)e + 1];
        off[(size_t)e] = (int32_t)a;
        tmp.clear();
        for (int64_t k = a; k < b; ++k) tmp.emplace_back(P.preds[(size_t)k].src, (int32_t)(k - a));
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const std::pair<int32_t, int32_t> &x, const std::pair<int32_t, int32_t> &y) {
                             return x.first < y.first;
                         });
        int32_t m = 0;
        int32_t lead = 0;
        for (size_t i = 0; i < tmp.size(); ++i) {
            if (i == 0 || tmp[i].first != tmp[i - 1].first) lead = tmp[i].second;  // the stable sort keeps list order
            first[(size_t)a + tmp[i].second] = lead;
            if (i > 0 && tmp[i].first == tmp[i - 1].first) continue;  // keep the first position
            key[(size_t)a + m] = tmp[i].first;
            pos[(size_t)a + m] = tmp[i].second;
            ++m;
        }
        nkey[(size_t)e] = m;
    }
    off[(size_t)P.E] = (int32_t)np;
    const size_t bytes = sizeof(int32_t) * (off.size() + key.size() + pos.size() + nkey.size() + first.size());
    void *blob = nullptr;
    hipError_t e = hipMalloc(&blob, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (predecessor table): ") + hipGetErrorString(e));
    auto *p = static_cast<int32_t *>(blob);
    int32_t *d_off = p, *d_key = d_off + off.size(), *d_pos = d_key + key.size(), *d_nkey = d_pos + pos.size();
    int32_t *d_first = d_nkey + nkey.size();
    e = hipMemcpy(d_off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && np) e = hipMemcpy(d_key, key.data(), sizeof(int32_t) * key.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && np) e = hipMemcpy(d_pos, pos.data(), sizeof(int32_t) * pos.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && P.E) e = hipMemcpy(d_nkey, nkey.data(), sizeof(int32_t) * nkey.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && np) e = hipMemcpy(d_first, first.data(), sizeof(int32_t) * first.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(blob);
        return hip_fail(e, "upload predecessor table");
    }
    d.topo = blob;
    d.pred_off = d_off;
    d.pred_key = d_key;
    d.pred_pos = d_pos;
    d.pred_nkey = d_nkey;
    d.pred_first = d_first;
    return MCDP_OK;
}

// numpy.quantile(..., method="linear") index arithmetic for n values
// (numpy/lib/_function_base_impl.py: _QuantileMethods["linear"], _get_indexes,
// _get_gamma), in double as numpy evaluates it.
int quantile_args(int64_t n, int32_t n_q, const double *q, QuantileArgs &qa) {
    if (n_q < 0 || n_q > kMaxQuantiles) return fail(MCDP_E_INVALID, "between 0 and 8 quantiles per call");
    if (n_q > 0 && !q) return fail(MCDP_E_INVALID, "quantiles is null");
    qa = QuantileArgs{};
    qa.n_q = n_q;
    auto rank_index = [&](int64_t r) {
        for (int i = 0; i < qa.n_ranks; ++i)
            if (qa.rank[i] == r) return i;
        qa.rank[qa.n_ranks] = (int32_t)r;
        return qa.n_ranks++;
    };
    for (int j = 0; j < n_q; ++j) {
        if (!(q[j] >= 0.0 && q[j] <= 1.0)) return fail(MCDP_E_INVALID, "quantiles must be in the range [0, 1]");
        const double v = (double)(n - 1) * q[j];
        double prev_index;  // numpy's previous_indexes (-1: the last value)
        int64_t lo, hi;
        if (v >= (double)(n - 1)) {
            prev_index = -1.0;
            lo = hi = n - 1;
        } else if (v < 0.0) {
            prev_index = 0.0;
            lo = hi = 0;
        } else {
            prev_index = std::floor(v);
            lo = (int64_t)prev_index;
            hi = lo + 1;
        }
        qa.lo[j] = rank_index(lo);
        qa.hi[j] = rank_index(hi);
        qa.gamma[j] = v - prev_index;
    }
    return MCDP_OK;
}

}  // namespace

int mcdp::chunk_stats_accumulate(mcdp_propagator &P, const double *realized, const int32_t *cause,
                                 const uint8_t *cause_pos, int32_t m, int32_t ld, bool first, ChunkStat *st,
                                 uint32_t *hist, int64_t *counts, hipStream_t s) {
    std::lock_guard<std::mutex> lock(P.mu);
    DeviceCopy *d = nullptr;
    int rc = ensure_device(P, d);
    if (rc) return rc;
    rc = ensure_topo(P, *d);
    if (rc) return rc;
    HIP_TRY(launch_chunk_accumulate(realized, m, ld, P.E, d->base.earliest, P.max_delay, kMaxChunkBins, first ? 1 : 0,
                                    st, hist, s),
            "chunk statistics kernel");
    if (cause_pos)
        HIP_TRY(launch_cause_pos_counts(cause_pos, m, ld, P.E, d->pred_off, d->pred_first, counts, first ? 0 : 1, s),
                "cause count kernel");
    else
        HIP_TRY(launch_cause_counts(cause, m, ld, P.E, d->pred_off, d->pred_key, d->pred_pos, d->pred_nkey, counts,
                                    first ? 0 : 1, s),
                "cause count kernel");
    return MCDP_OK;
}

int mcdp::chunk_stats_finish(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                             const uint32_t *hist, double *mean, double *quantiles, hipStream_t s) {
    QuantileArgs qa;
    int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(P.mu);
    DeviceCopy *d = nullptr;
    rc = ensure_device(P, d);
    if (rc) return rc;
    HIP_TRY(launch_chunk_finish(n, P.E, qa, d->base.earliest, P.max_delay, kMaxChunkBins, st, hist, mean, quantiles, s),
            "chunk statistics kernel");
    return MCDP_OK;
}

int mcdp::exact_rank_count(int64_t n, int32_t n_q, const double *q, int32_t *n_ranks) {
    QuantileArgs qa;
    const int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    *n_ranks = qa.n_ranks;
    return MCDP_OK;
}

int mcdp::exact_stats_init(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                           const uint32_t *hist, ExactWork &w, const WatchWork *watch, hipStream_t s) {
    QuantileArgs qa;
    int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    if (qa.n_ranks > w.n_ranks) return fail(MCDP_E_INVALID, "internal error: exact quantile workspace too small");
    w.n_ranks = qa.n_ranks;
    std::lock_guard<std::mutex> lock(P.mu);
    DeviceCopy *d = nullptr;
    rc = ensure_device(P, d);
    if (rc) return rc;
    HIP_TRY(launch_exact_init(n, P.E, qa, st, hist, kMaxChunkBins, d->base.earliest, P.max_delay, w, watch, s),
            "exact quantile kernels");
    return MCDP_OK;
}

int mcdp::counters_clear(uint64_t *c, int n, hipStream_t s) {
    HIP_TRY(launch_counters_clear(c, n, s), "counter clearing kernel");
    return MCDP_OK;
}

int mcdp::watch_select(mcdp_propagator &P, int64_t n, int32_t m, int32_t n_q, const double *q, const ChunkStat *st,
                       const double *ranges, const WatchWork &w, hipStream_t s) {
    QuantileArgs qa;
    const int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    HIP_TRY(launch_watch_select(n, m, P.E, qa, st, ranges, w, s), "kept-row selection kernel");
    return MCDP_OK;
}

int mcdp::watch_copy(const double *realized, int32_t m, int32_t ld, int64_t c0, const WatchWork &w, hipStream_t s) {
    HIP_TRY(launch_watch_copy(realized, m, ld, c0, w, s), "kept-row copy kernel");
    return MCDP_OK;
}

int mcdp::watch_finish(int64_t n, int32_t n_q, const double *q, const WatchWork &w, int32_t rows, double *quantiles,
                       hipStream_t s) {
    if (rows <= 0 || n_q <= 0) return MCDP_OK;
    if (rows > w.cap || n > std::numeric_limits<int32_t>::max() || w.ld != n)
        return fail(MCDP_E_INVALID, "internal error: kept rows do not match the request");
    QuantileArgs qa;
    const int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    HIP_TRY(launch_event_quantiles(w.store, (int32_t)n, (int32_t)n, rows, qa, w.wq, nullptr, s), "quantile kernel");
    HIP_TRY(launch_watch_scatter(w, rows, n_q, quantiles, s), "kept-row scatter kernel");
    return MCDP_OK;
}

int mcdp::exact_stats_pass(mcdp_propagator &P, const double *realized, int32_t m, int32_t ld, const ExactWork &w,
                           hipStream_t s) {
    std::lock_guard<std::mutex> lock(P.mu);
    DeviceCopy *d = nullptr;
    int rc = ensure_device(P, d);
    if (rc) return rc;
    HIP_TRY(launch_exact_pass(realized, m, ld, P.E, d->base.earliest, P.max_delay, kMaxChunkBins, w, s),
            "exact quantile pass kernel");
    return MCDP_OK;
}

int mcdp::exact_stats_round_end(mcdp_propagator &P, const ExactWork &w, hipStream_t s) {
    HIP_TRY(launch_exact_round_end(P.E, w, s), "exact quantile kernels");
    return MCDP_OK;
}

int mcdp::exact_stats_finish(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                             const ExactWork &w, double *quantiles, hipStream_t s) {
    QuantileArgs qa;
    const int rc = quantile_args(n, n_q, q, qa);
    if (rc) return rc;
    HIP_TRY(launch_exact_finish(P.E, qa, st, w, quantiles, s), "exact quantile kernel");
    return MCDP_OK;
}

int mcdp::row_set(mcdp_propagator &P, const double *realized, const int32_t *cause, const uint8_t *cause_pos,
                  int64_t n, int32_t n_q, const double *q, double *mean, double *quantiles, int64_t *counts,
                  int64_t row0, RowSet &rs) {
    if (n < 1 || n > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "statistics need between 1 and 2^31 - 1 samples");
    if (row0 < 0 || row0 + P.E > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "too many events for one statistics launch");
    if (!realized || !mean || (n_q > 0 && !quantiles) || (counts && !cause && !cause_pos))
        return fail(MCDP_E_INVALID, "null argument");
    rs = RowSet{};
    int rc = quantile_args(n, n_q, q, rs.qa);
    if (rc) return rc;
    rs.realized = realized;
    rs.cause = cause_pos ? nullptr : cause;
    rs.cause_pos = cause_pos;
    rs.mean = mean;
    rs.quant = quantiles;
    rs.counts = counts;
    rs.n = (int32_t)n;
    rs.ld = (uint32_t)n;
    rs.E = P.E;
    rs.row0 = (int32_t)row0;
    if (counts) {
        std::lock_guard<std::mutex> lock(P.mu);
        DeviceCopy *d = nullptr;
        rc = ensure_device(P, d);
        if (rc) return rc;
        rc = ensure_topo(P, *d);
        if (rc) return rc;
        rs.pred_off = d->pred_off;
        rs.pred_key = d->pred_key;
        rs.pred_pos = d->pred_pos;
        rs.pred_nkey = d->pred_nkey;
        rs.pred_first = d->pred_first;
    }
    return MCDP_OK;
}

int mcdp::set_stats(const RowSet *sets_dev, int32_t n_sets, int64_t rows, bool quantiles, bool counts,
                    hipStream_t s) {
    if (rows > std::numeric_limits<int32_t>::max()) return fail(MCDP_E_INVALID, "too many rows for one launch");
    HIP_TRY(launch_set_stats(sets_dev, n_sets, (int32_t)rows, quantiles, counts, s), "statistics kernels");
    return MCDP_OK;
}

int mcdp::event_stats_positions(mcdp_propagator &P, const double *realized, const uint8_t *cause_pos, int64_t n,
                                int64_t ld, int32_t n_q, const double *q, double *mean, double *quantiles,
                                int64_t *counts, hipStream_t s) {
    if (n < 1 || ld < n || ld > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "statistics need 1 <= n <= ld < 2^31");
    if (!realized || (counts && !cause_pos)) return fail(MCDP_E_INVALID, "null argument");
    QuantileArgs qa;
    int rc = quantile_args(n, quantiles ? n_q : 0, q, qa);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(P.mu);
    if (quantiles && qa.n_q > 0)
        HIP_TRY(launch_event_quantiles(realized, (int32_t)n, (int32_t)ld, P.E, qa, quantiles, mean, s),
                "quantile kernel");
    else if (mean)
        HIP_TRY(launch_event_mean(realized, (int32_t)n, (int32_t)ld, P.E, mean, s), "mean kernel");
    if (counts) {
        DeviceCopy *d = nullptr;
        rc = ensure_device(P, d);
        if (rc) return rc;
        rc = ensure_topo(P, *d);
        if (rc) return rc;
        HIP_TRY(launch_cause_pos_counts(cause_pos, (int32_t)n, (int32_t)ld, P.E, d->pred_off, d->pred_first, counts, 0,
                                        s),
                "cause count kernel");
    }
    return MCDP_OK;
}

int mcdp::prepare_stats(mcdp_propagator &P) {
    std::lock_guard<std::mutex> lock(P.mu);
    DeviceCopy *d = nullptr;
    int rc = ensure_device(P, d);
    if (rc) return rc;
    return ensure_topo(P, *d);
}

extern "C" {

int mcdp_propagator_pred_offsets(const mcdp_propagator *prop, int64_t *out) {
    if (!prop || !out) return fail(MCDP_E_INVALID, "null argument");
    std::copy(prop->pred_off.begin(), prop->pred_off.end(), out);
    return MCDP_OK;
}

int mcdp_event_stats_device(mcdp_propagator *prop, const double *realized_dev, const int32_t *cause_dev, int64_t n,
                            int64_t ld, int32_t n_q, const double *q, double *mean_dev, double *quantiles_dev,
                            int64_t *cause_counts_dev, void *stream) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n < 1) return fail(MCDP_E_INVALID, "statistics need at least one sample");
    if (ld < n) return fail(MCDP_E_INVALID, "ld must be >= n");
    if (n > std::numeric_limits<int32_t>::max() || ld > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "n and ld must be below 2^31");
    if ((mean_dev || quantiles_dev) && !realized_dev) return fail(MCDP_E_INVALID, "realized is null");
    if (cause_counts_dev && !cause_dev) return fail(MCDP_E_INVALID, "cause is null");
    QuantileArgs qa;
    int rc = quantile_args(n, quantiles_dev ? n_q : 0, q, qa);
    if (rc) return rc;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: statistics run on the GPU only");
    std::lock_guard<std::mutex> lock(prop->mu);
    const auto s = static_cast<hipStream_t>(stream);
    const int32_t E = prop->E;
    if (quantiles_dev && qa.n_q > 0)
        HIP_TRY(launch_event_quantiles(realized_dev, (int32_t)n, (int32_t)ld, E, qa, quantiles_dev, mean_dev, s),
                "quantile kernel");
    else if (mean_dev)
        HIP_TRY(launch_event_mean(realized_dev, (int32_t)n, (int32_t)ld, E, mean_dev, s), "mean kernel");
    if (cause_counts_dev) {
        DeviceCopy *d = nullptr;
        rc = ensure_device(*prop, d);
        if (rc) return rc;
        rc = ensure_topo(*prop, *d);
        if (rc) return rc;
        HIP_TRY(launch_cause_counts(cause_dev, (int32_t)n, (int32_t)ld, E, d->pred_off, d->pred_key, d->pred_pos,
                                    d->pred_nkey, cause_counts_dev, 0, s),
                "cause count kernel");
    }
    return MCDP_OK;
}

int mcdp_run_stats_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, int32_t n_q, const double *q,
                        double *mean, double *quantiles, int64_t *cause_counts) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n < 1 || !seeds) return fail(MCDP_E_INVALID, "statistics need at least one seed");
    const int64_t E = prop->E, P = (int64_t)prop->preds.size();
    if (E == 0) return MCDP_OK;
    QuantileArgs qa;
    int rc = quantile_args(n, quantiles ? n_q : 0, q, qa);
    if (rc) return rc;
    // causes as one-byte winner positions when every event has at most 254
    // entries (the same counts, a quarter of the bytes)
    const bool pos = max_entries(*prop) <= kMaxPositionEntries;
    const size_t cb = pos ? 1 : 4;
    const size_t o_seeds = 0, o_real = 256 * (((size_t)n * 4 + 255) / 256);
    const size_t o_cause = o_real + (size_t)(E * n) * 8;
    const size_t o_mean = o_cause + 256 * (((size_t)(E * n) * cb + 255) / 256);
    const size_t o_q = o_mean + (size_t)E * 8;
    const size_t o_cnt = o_q + (size_t)E * (size_t)std::max(n_q, 0) * 8;
    const size_t bytes = o_cnt + (size_t)(E + P) * 8;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: statistics run on the GPU only");
    std::lock_guard<std::mutex> host_lock(prop->host_mu);
    void *buf = nullptr;
    hipStream_t st = nullptr;
    rc = host_workspace(*prop, bytes, buf, st);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    auto *b = static_cast<unsigned char *>(buf);
    auto *d_seeds = reinterpret_cast<int32_t *>(b + o_seeds);
    auto *d_real = reinterpret_cast<double *>(b + o_real);
    auto *d_cause = reinterpret_cast<int32_t *>(b + o_cause);
    auto *d_mean = reinterpret_cast<double *>(b + o_mean);
    auto *d_q = reinterpret_cast<double *>(b + o_q);
    auto *d_cnt = reinterpret_cast<int64_t *>(b + o_cnt);
    e = hipMemcpyAsync(d_seeds, seeds, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy");
    auto *d_pos = reinterpret_cast<uint8_t *>(d_cause);
    if (!rc && pos && cause_counts) {
        rc = run_many_positions(*prop, d_seeds, n, n, d_real, d_pos, st);
        if (!rc)
            rc = event_stats_positions(*prop, d_real, d_pos, n, n, n_q, q, mean ? d_mean : nullptr,
                                       quantiles ? d_q : nullptr, d_cnt, st);
    } else {
        if (!rc)
            rc = mcdp_run_many_device(prop, d_seeds, n, n, d_real, nullptr, cause_counts ? d_cause : nullptr, st);
        if (!rc)
            rc = mcdp_event_stats_device(prop, d_real, d_cause, n, n, n_q, q, mean ? d_mean : nullptr,
                                         quantiles ? d_q : nullptr, cause_counts ? d_cnt : nullptr, st);
    }
    if (!rc && mean &&
        (e = hipMemcpyAsync(mean, d_mean, (size_t)E * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy");
    if (!rc && quantiles && n_q > 0 &&
        (e = hipMemcpyAsync(quantiles, d_q, (size_t)E * (size_t)n_q * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy");
    if (!rc && cause_counts &&
        (e = hipMemcpyAsync(cause_counts, d_cnt, (size_t)(E + P) * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess && !rc) rc = hip_fail(e, "statistics");
    host_workspace_trim(*prop);
    return rc;
}

}  // extern "C"
