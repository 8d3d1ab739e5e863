// Serving: per-event statistics for a stream of requests over resident
// compiled contexts (C ABI in include/mcdp.h).  The reference answers one
// run_many call at a time on the calling thread (_core.cpp:355-361) and leaves
// statistics to its callers; here a scheduler thread owns the device:
//   * contexts are uploaded once (plan + predecessor table) and stay in HBM;
//   * every request reserves its footprint (seeds, realized and cause rows,
//     statistics) from one device pool of a fixed budget; requests are
//     admitted in arrival order as soon as the pool and a running slot allow,
//     also while others are on the device (no batch boundary);
//   * an admitted request runs run_many (no durations output) and the
//     statistics kernels on one of a few streams, so requests overlap on the
//     device, and copies its results into pinned host memory;
//   * completion is seen through an event, which returns the memory at once.
// A request's results are those of mcdp_run_stats_host on the same seeds, bit
// for bit: the kernels do not depend on what else runs.  A request larger than
// the per-request cap runs in chunks of samples: running mean, extremes, cause
// counts and a 4,096-bin histogram per event over the chunks, then further
// passes over the same chunks (their results are a function of the seeds) that
// gather or narrow down the values around each wanted order statistic, so its
// quantiles are exact too.  Between passes the scheduler reads a few counters
// back and launches the next pass; other requests keep running meanwhile.
// Most events' order statistics lie on the masses of their extremes (a cap
// that binds, an earliest time that holds); the events whose first chunk
// says otherwise keep their whole rows across the chunks, so their quantiles
// need no further pass either, and a request usually needs none.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"

using namespace mcdp;

namespace {

using Clock = std::chrono::steady_clock;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Exact quantiles of chunked requests.  A target's histogram bin holds about
// n / 4,096 values where the row's values spread evenly, so candidate space is
// n / 512 values per target on average (at least 16, at most 256; one target
// may take up to kExactMaxCollect of it); refine slots: one per 64 targets
// (at least 16, at most 1,024).  A request may take kMaxExactRounds further
// passes before it fails (a refining pass shrinks a region 256-fold, so a few
// suffice).
constexpr int64_t kExactCandPerTarget = 256;
constexpr int32_t kExactRefineSlots = 1024;
constexpr int32_t kMaxExactRounds = 64;

struct ExactSizes {
    int64_t targets = 0, cand_per_target = 0;
    int32_t slots = 0;
    size_t bytes = 0;
};
// `kept`: rows kept whole (WatchWork), each with its n samples and quantiles.
ExactSizes exact_sizes(int64_t n, int64_t E, int64_t n_q, int64_t kept) {
    ExactSizes z;
    if (n_q <= 0) return z;
    z.targets = E * std::min<int64_t>(2 * n_q, kMaxRanks);
    z.cand_per_target = std::min<int64_t>(kExactCandPerTarget, std::max<int64_t>(16, (n + 511) / 512));
    z.slots = (int32_t)std::min<int64_t>(kExactRefineSlots, std::max<int64_t>(16, (z.targets + 63) / 64));
    z.bytes = align_up((size_t)z.targets * sizeof(ExactTarget), 256) +
              align_up((size_t)(z.targets * z.cand_per_target) * 8, 256) +
              align_up((size_t)z.slots * kExactSubBins * 4, 256) + 2 * align_up((size_t)z.slots * kExactSubBins * 8, 256) +
              256;
    if (kept > 0)
        z.bytes += align_up((size_t)kept * 4, 256) + align_up((size_t)kept * (size_t)n * 8, 256) +
                   align_up((size_t)(kept * n_q) * 8, 256);
    return z;
}

// Rows kept whole for a chunked request: one per 128 events (at least 16, at
// most E; `limit` >= 0 replaces that), within a quarter of the request cap.
// Events whose first chunk puts a rank outside the extremes' masses are
// rare where caps bind (a few per thousand events), and those past the
// limit take further passes instead.
int64_t kept_rows(int64_t n, int64_t E, int64_t n_q, size_t cap, int32_t limit) {
    if (n_q <= 0 || E <= 0 || n <= 0) return 0;
    int64_t k = limit >= 0 ? std::min<int64_t>(limit, E) : std::min<int64_t>(E, std::max<int64_t>(16, (E + 127) / 128));
    return std::min<int64_t>(k, (int64_t)(cap / 4 / ((size_t)n * 8)));
}

// Best-fit allocator over [0, size) with coalescing free blocks.
class Pool {
   public:
    explicit Pool(size_t size) : size_(size) {
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t &off) {
        bytes = align_up(std::max<size_t>(bytes, 1), 256);
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return false;
        off = best->first;
        const size_t rest = best->second - bytes;
        free_.erase(best);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1), 256);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t size() const { return size_; }

   private:
    size_t size_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

enum class State { Queued, Running, Done, Failed };

struct Request {
    int64_t id = 0;
    int32_t ctx = 0;
    std::vector<int32_t> seeds;
    std::vector<double> q;
    int64_t E = 0, P = 0;
    int64_t chunk = 0;  // > 0: run in chunks of this many samples (statistics over chunks)
    size_t bytes = 0, off = 0;
    int stream = 0;
    hipEvent_t done = nullptr;
    double *h_out = nullptr;  // pinned: mean [E], quantiles [E * n_q], cause counts [E + P] (as int64)
    // device buffers (in the pool) and, for chunked requests with quantiles,
    // the exact selection: phase 1 while a pass (or its set-up) runs, whose
    // counters come back to h_status; phase 2 once the results are on the way
    int32_t *d_seeds = nullptr, *d_cause = nullptr;  // d_cause holds winner positions when cause_bytes == 1
    double *d_real = nullptr, *d_mean = nullptr, *d_quant = nullptr;
    int64_t *d_cnt = nullptr;
    ChunkStat *d_st = nullptr;
    uint32_t *d_hist = nullptr;
    ExactWork exact{};
    int64_t kept_cap = 0;  // rows kept whole (0: none)
    int cause_bytes = 4;   // 1 when the walks write winner positions
    WatchWork watch{};
    int phase = 0;
    int rounds = 0;
    int64_t buf_c0 = 0;  // first sample of the chunk whose rows d_real holds
    uint64_t *h_status = nullptr;  // pinned: the exact selection's counters and the kept-row count
    Clock::time_point t_submit, t_start, t_end;
    State state = State::Queued;
    int rc = MCDP_OK;
    std::string error;
};

// Device bytes of a request: seeds, realized and cause rows for `rows`
// samples at a time (cb bytes per cause: 4 for cause events, 1 for winner
// positions), statistics (and, run in chunks, their running state, the exact
// selection's workspace and `kept` rows kept whole).
size_t footprint(int64_t n, int64_t rows, int64_t E, int64_t P, int64_t n_q, bool chunked, int64_t kept, int cb) {
    size_t b = align_up((size_t)n * 4, 256) + align_up((size_t)(E * rows) * 8, 256) +
               align_up((size_t)(E * rows) * (size_t)cb, 256) + align_up((size_t)E * 8, 256) +
               align_up((size_t)(E * n_q) * 8, 256) + align_up((size_t)(E + P) * 8, 256);
    if (chunked) b += align_up((size_t)E * sizeof(ChunkStat), 256) + align_up((size_t)E * kMaxChunkBins * 4, 256);
    if (chunked) b += exact_sizes(n, E, n_q, kept).bytes;  // targets, candidates, refine slots, counters
    return b;
}

// Samples per chunk of a request of n samples whose whole footprint exceeds
// `cap`: the largest multiple of 128 whose footprint (with the chunks' running
// statistics) fits the cap, at most n; 0 if not even 128 samples fit.  The
// 1,024 bytes of slack cover the 256-byte alignment of the two per-sample
// row blocks.
int64_t chunk_samples(int64_t n, int64_t E, int64_t P, int64_t n_q, size_t cap, int64_t kept, int cb) {
    const size_t fixed = footprint(n, 0, E, P, n_q, true, kept, cb);
    const size_t per_sample = (size_t)std::max<int64_t>(E, 1) * (size_t)(8 + cb);
    if (fixed >= cap || cap - fixed <= 1024) return 0;
    int64_t m = (int64_t)((cap - fixed - 1024) / per_sample);
    m = m / 128 * 128;
    if (m < 128) return 0;
    m = std::min<int64_t>(m, n);
    return footprint(n, m, E, P, n_q, true, kept, cb) <= cap ? m : 0;
}

// Cause bytes per event and sample of a request: its walks write the winning
// entry's position (one byte) when no event has more than 254 entries, the
// cause event (four) otherwise.  The counts come out the same.
int cause_bytes_for(int32_t max_entries) { return max_entries <= kMaxPositionEntries ? 1 : 4; }

// A chunked request's samples per chunk and rows kept whole: the kept rows
// give way (to none) when the chunks would not fit beside them.
void chunk_plan(int64_t n, int64_t E, int64_t P, int64_t n_q, size_t cap, bool exact, int32_t limit, int cb,
                int64_t &m, int64_t &kept) {
    kept = exact ? kept_rows(n, E, n_q, cap, limit) : 0;
    m = chunk_samples(n, E, P, n_q, cap, kept, cb);
    if (m < 1 && kept > 0) {
        kept = 0;
        m = chunk_samples(n, E, P, n_q, cap, 0, cb);
    }
}

}  // namespace

struct mcdp_server {
    int device = 0;
    void *pool_base = nullptr;
    std::unique_ptr<Pool> pool;
    int32_t max_running = 4;
    size_t request_cap = 0;  // a request needing more runs in chunks
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;
    std::vector<mcdp_propagator *> contexts;
    std::vector<double *> ranges;  // per context: [E][2] bounds of realized times (null: not finite)
    std::vector<int32_t> max_entries;  // per context: the most predecessor entries of an event

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queued;
    std::vector<Request *> running;
    std::unordered_map<int64_t, std::unique_ptr<Request>> all;
    int64_t next_id = 1;
    int64_t n_submitted = 0, n_completed = 0, n_failed = 0, max_running_seen = 0, n_chunked = 0, n_rounds = 0,
            n_kept = 0;
    bool stop = false;
    bool exact_chunked = true;  // exact quantiles for chunked requests (further passes)
    int32_t exact_c_max = kExactMaxCollect, exact_slots = kExactRefineSlots;
    int64_t exact_cand_per_target = kExactCandPerTarget;
    int32_t kept_limit = -1;  // rows kept whole per chunked request: -1 by kept_rows, else at most this many
    std::thread worker;

    void run();
    void start(Request &r);
    void finish_failed(Request &r);
};

namespace {

int copy_results(mcdp_server &srv, Request &r);

// Launches a request's work on its stream; the caller holds srv.mu.
int launch_request(mcdp_server &srv, Request &r) {
    mcdp_propagator *prop = srv.contexts[(size_t)r.ctx];
    const int64_t n = (int64_t)r.seeds.size(), E = r.E, nq = (int64_t)r.q.size();
    const int64_t rows = r.chunk > 0 ? r.chunk : n;
    auto *base = static_cast<unsigned char *>(srv.pool_base) + r.off;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + o;
        o += align_up(bytes, 256);
        return p;
    };
    auto *d_seeds = r.d_seeds = reinterpret_cast<int32_t *>(take((size_t)n * 4));
    auto *d_real = r.d_real = reinterpret_cast<double *>(take((size_t)(E * rows) * 8));
    auto *d_cause = r.d_cause = reinterpret_cast<int32_t *>(take((size_t)(E * rows) * (size_t)r.cause_bytes));
    auto *d_pos = r.cause_bytes == 1 ? reinterpret_cast<uint8_t *>(d_cause) : nullptr;
    auto *d_mean = r.d_mean = reinterpret_cast<double *>(take((size_t)E * 8));
    auto *d_quant = r.d_quant = reinterpret_cast<double *>(take((size_t)(E * nq) * 8));
    auto *d_cnt = r.d_cnt = reinterpret_cast<int64_t *>(take((size_t)(E + r.P) * 8));
    const hipStream_t s = srv.streams[(size_t)r.stream];
    HIP_TRY(hipMemcpyAsync(d_seeds, r.seeds.data(), (size_t)n * 4, hipMemcpyHostToDevice, s), "copy seeds");
    int rc;
    if (r.chunk > 0) {
        // run_many chunk by chunk into the same rows; statistics merged across chunks
        auto *d_st = r.d_st = reinterpret_cast<ChunkStat *>(take((size_t)E * sizeof(ChunkStat)));
        auto *d_hist = r.d_hist = reinterpret_cast<uint32_t *>(take((size_t)E * kMaxChunkBins * 4));
        const bool exact = srv.exact_chunked && nq > 0;
        ExactWork &w = r.exact;
        WatchWork &v = r.watch;
        if (exact) {
            // the exact selection's workspace, and the rows kept whole
            const ExactSizes z = exact_sizes(n, E, nq, r.kept_cap);
            w.targets = reinterpret_cast<ExactTarget *>(take((size_t)z.targets * sizeof(ExactTarget)));
            w.cand = reinterpret_cast<double *>(take((size_t)(z.targets * z.cand_per_target) * 8));
            w.slot_cnt = reinterpret_cast<uint32_t *>(take((size_t)z.slots * kExactSubBins * 4));
            w.slot_min = reinterpret_cast<uint64_t *>(take((size_t)z.slots * kExactSubBins * 8));
            w.slot_max = reinterpret_cast<uint64_t *>(take((size_t)z.slots * kExactSubBins * 8));
            w.counters = reinterpret_cast<uint64_t *>(take(256));
            w.n_ranks = (int32_t)std::min<int64_t>(2 * nq, kMaxRanks);
            w.n_slots = std::min(srv.exact_slots, z.slots);
            w.c_max = std::min(srv.exact_c_max, kExactMaxCollect);
            w.cand_cap = (uint32_t)std::min<int64_t>(z.targets * std::min(srv.exact_cand_per_target, z.cand_per_target),
                                                     std::numeric_limits<int32_t>::max());
            v.count = w.counters + 4;
            v.ld = n;
            v.cap = 0;
            if (r.kept_cap > 0) {
                v.list = reinterpret_cast<int32_t *>(take((size_t)r.kept_cap * 4));
                v.store = reinterpret_cast<double *>(take((size_t)r.kept_cap * (size_t)n * 8));
                v.wq = reinterpret_cast<double *>(take((size_t)(r.kept_cap * nq) * 8));
                if (srv.ranges[(size_t)r.ctx]) v.cap = (int32_t)r.kept_cap;
            }
            if (o > r.bytes) return fail(MCDP_E_INVALID, "internal error: request layout exceeds its footprint");
            rc = counters_clear(w.counters, 5, s);
            if (rc) return rc;
        }
        for (int64_t c0 = 0; c0 < n; c0 += r.chunk) {
            const int64_t m = std::min<int64_t>(r.chunk, n - c0);
            rc = d_pos ? run_many_positions(*prop, d_seeds + c0, m, r.chunk, d_real, d_pos, s)
                       : mcdp_run_many_device(prop, d_seeds + c0, m, r.chunk, d_real, nullptr, d_cause, s);
            if (rc) return rc;
            rc = chunk_stats_accumulate(*prop, d_real, d_pos ? nullptr : d_cause, d_pos, (int32_t)m, (int32_t)r.chunk,
                                        c0 == 0, d_st, d_hist, d_cnt, s);
            if (rc) return rc;
            if (v.cap > 0) {
                if (c0 == 0) {
                    rc = watch_select(*prop, n, (int32_t)m, (int32_t)nq, r.q.data(), d_st, srv.ranges[(size_t)r.ctx],
                                      v, s);
                    if (rc) return rc;
                }
                rc = watch_copy(d_real, (int32_t)m, (int32_t)r.chunk, c0, v, s);
                if (rc) return rc;
            }
            r.buf_c0 = c0;
        }
        rc = chunk_stats_finish(*prop, n, (int32_t)nq, r.q.data(), d_st, d_hist, d_mean, exact ? nullptr : d_quant, s);
        if (rc) return rc;
        if (exact) {
            // targets from the histograms (those of kept rows resolved); the
            // scheduler continues once the counters are back
            rc = exact_stats_init(*prop, n, (int32_t)nq, r.q.data(), d_st, d_hist, w, v.cap > 0 ? &v : nullptr, s);
            if (rc) return rc;
            HIP_TRY(hipMemcpyAsync(r.h_status, w.counters, 5 * 8, hipMemcpyDeviceToHost, s), "copy counters");
            HIP_TRY(hipEventCreateWithFlags(&r.done, hipEventDisableTiming), "hipEventCreate");
            HIP_TRY(hipEventRecord(r.done, s), "hipEventRecord");
            r.phase = 1;
            return MCDP_OK;
        }
    } else if (d_pos) {
        rc = run_many_positions(*prop, d_seeds, n, n, d_real, d_pos, s);
        if (rc) return rc;
        rc = event_stats_positions(*prop, d_real, d_pos, n, n, (int32_t)nq, r.q.data(), d_mean, nq ? d_quant : nullptr,
                                   d_cnt, s);
    } else {
        rc = mcdp_run_many_device(prop, d_seeds, n, n, d_real, nullptr, d_cause, s);
        if (rc) return rc;
        rc = mcdp_event_stats_device(prop, d_real, d_cause, n, n, (int32_t)nq, r.q.data(), d_mean,
                                     nq ? d_quant : nullptr, d_cnt, s);
    }
    if (rc) return rc;
    return copy_results(srv, r);
}

// Enqueues the copies of a request's results and its completion event.
int copy_results(mcdp_server &srv, Request &r) {
    const int64_t E = r.E, nq = (int64_t)r.q.size();
    const hipStream_t s = srv.streams[(size_t)r.stream];
    double *h = r.h_out;
    HIP_TRY(hipMemcpyAsync(h, r.d_mean, (size_t)E * 8, hipMemcpyDeviceToHost, s), "copy results");
    if (nq) HIP_TRY(hipMemcpyAsync(h + E, r.d_quant, (size_t)(E * nq) * 8, hipMemcpyDeviceToHost, s), "copy results");
    HIP_TRY(hipMemcpyAsync(h + E + E * nq, r.d_cnt, (size_t)(E + r.P) * 8, hipMemcpyDeviceToHost, s), "copy results");
    HIP_TRY(hipEventCreateWithFlags(&r.done, hipEventDisableTiming), "hipEventCreate");
    HIP_TRY(hipEventRecord(r.done, s), "hipEventRecord");
    r.phase = 2;
    return MCDP_OK;
}

// The exact selection's next step, once the counters of the last one are
// back: another pass over every chunk while targets are open, else the
// quantiles and the results' copies.
int continue_exact(mcdp_server &srv, Request &r) {
    const uint64_t open = r.h_status[2], bad = r.h_status[3];
    if (bad) return fail(MCDP_E_DEVICE, "internal error: exact quantile selection found inconsistent counts");
    mcdp_propagator *prop = srv.contexts[(size_t)r.ctx];
    const hipStream_t s = srv.streams[(size_t)r.stream];
    const int64_t n = (int64_t)r.seeds.size(), nq = (int64_t)r.q.size();
    if (open == 0) {
        int rc = exact_stats_finish(*prop, n, (int32_t)nq, r.q.data(), r.d_st, r.exact, r.d_quant, s);
        if (rc) return rc;
        // the kept rows' quantiles, from their whole rows, over those
        const int32_t rows = (int32_t)std::min<uint64_t>(r.h_status[4], (uint64_t)r.watch.cap);
        rc = watch_finish(n, (int32_t)nq, r.q.data(), r.watch, rows, r.d_quant, s);
        if (rc) return rc;
        srv.n_kept += rows;
        return copy_results(srv, r);
    }
    if (r.rounds >= kMaxExactRounds) return fail(MCDP_E_DEVICE, "exact quantile selection did not converge");
    ++r.rounds;
    ++srv.n_rounds;
    // the chunk still in the rows first, then the others walked again (no cause rows: the counts are done)
    const int64_t kept = r.buf_c0;
    int rc = exact_stats_pass(*prop, r.d_real, (int32_t)std::min<int64_t>(r.chunk, n - kept), (int32_t)r.chunk,
                              r.exact, s);
    if (rc) return rc;
    for (int64_t c0 = 0; c0 < n; c0 += r.chunk) {
        if (c0 == kept) continue;
        const int64_t m = std::min<int64_t>(r.chunk, n - c0);
        rc = mcdp_run_many_device(prop, r.d_seeds + c0, m, r.chunk, r.d_real, nullptr, nullptr, s);
        if (rc) return rc;
        rc = exact_stats_pass(*prop, r.d_real, (int32_t)m, (int32_t)r.chunk, r.exact, s);
        if (rc) return rc;
        r.buf_c0 = c0;
    }
    rc = exact_stats_round_end(*prop, r.exact, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(r.h_status, r.exact.counters, 5 * 8, hipMemcpyDeviceToHost, s), "copy counters");
    HIP_TRY(hipEventCreateWithFlags(&r.done, hipEventDisableTiming), "hipEventCreate");
    HIP_TRY(hipEventRecord(r.done, s), "hipEventRecord");
    r.phase = 1;
    return MCDP_OK;
}

}  // namespace

void mcdp_server::start(Request &r) {
    int best = 0;
    for (size_t i = 1; i < streams.size(); ++i)
        if (stream_load[i] < stream_load[(size_t)best]) best = (int)i;
    r.stream = best;
    r.t_start = Clock::now();
    r.rc = launch_request(*this, r);
    if (r.rc) {
        finish_failed(r);
        return;
    }
    ++stream_load[(size_t)r.stream];
    r.state = State::Running;
    running.push_back(&r);
    max_running_seen = std::max<int64_t>(max_running_seen, (int64_t)running.size());
}

// A request whose launch failed: its enqueued work finishes before its memory
// returns to the pool.  The caller holds mu; r is not in `running`.
void mcdp_server::finish_failed(Request &r) {
    r.error = mcdp_last_error();
    r.state = State::Failed;
    (void)hipStreamSynchronize(streams[(size_t)r.stream]);
    if (r.done) {
        (void)hipEventDestroy(r.done);
        r.done = nullptr;
    }
    pool->release(r.off, r.bytes);
    r.t_end = Clock::now();
    ++n_failed;
    cv_done.notify_all();
}

void mcdp_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lock(mu);
    while (true) {
        // completions
        bool changed = false;
        for (size_t i = 0; i < running.size();) {
            Request &r = *running[i];
            const hipError_t e = hipEventQuery(r.done);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (e == hipSuccess && r.phase == 1) {  // an exact-selection step is done: launch the next
                (void)hipEventDestroy(r.done);
                r.done = nullptr;
                r.rc = continue_exact(*this, r);
                if (r.rc == MCDP_OK) {
                    ++i;
                    continue;
                }
                --stream_load[(size_t)r.stream];
                running.erase(running.begin() + (std::ptrdiff_t)i);
                finish_failed(r);
                changed = true;
                continue;
            }
            r.t_end = Clock::now();
            if (e == hipSuccess) {
                r.state = State::Done;
                ++n_completed;
            } else {
                r.state = State::Failed;
                r.rc = MCDP_E_DEVICE;
                r.error = std::string("request failed on the device: ") + hipGetErrorString(e);
                ++n_failed;
            }
            (void)hipEventDestroy(r.done);
            r.done = nullptr;
            pool->release(r.off, r.bytes);
            --stream_load[(size_t)r.stream];
            running.erase(running.begin() + (std::ptrdiff_t)i);
            changed = true;
        }
        if (changed) cv_done.notify_all();
        // admissions, in arrival order, while a slot and the memory allow
        while (!queued.empty() && (int32_t)running.size() < max_running) {
            Request &r = *queued.front();
            size_t off = 0;
            if (!pool->alloc(r.bytes, off)) break;
            r.off = off;
            queued.pop_front();
            start(r);
        }
        if (stop && running.empty() && queued.empty()) break;
        if (running.empty())
            cv_work.wait(lock, [&] { return stop || !queued.empty(); });
        else
            cv_work.wait_for(lock, std::chrono::microseconds(50));
    }
}

extern "C" {

int mcdp_server_create(int64_t memory_budget, int64_t request_cap, int32_t max_running, int32_t n_streams,
                       mcdp_server **out) {
    if (!out) return fail(MCDP_E_INVALID, "out is null");
    *out = nullptr;
    if (memory_budget <= 0) return fail(MCDP_E_INVALID, "memory_budget must be positive");
    if (request_cap < 0 || request_cap > memory_budget)
        return fail(MCDP_E_INVALID, "request_cap must be between 0 and memory_budget");
    if (max_running < 1 || n_streams < 1 || n_streams > 16)
        return fail(MCDP_E_INVALID, "need max_running >= 1 and 1 <= n_streams <= 16");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: serving runs on the GPU only");
    auto srv = std::make_unique<mcdp_server>();
    HIP_TRY(hipGetDevice(&srv->device), "hipGetDevice");
    hipError_t e = hipMalloc(&srv->pool_base, (size_t)memory_budget);
    if (e != hipSuccess)
        return fail(MCDP_E_NOMEM, "hipMalloc (serving pool of " + std::to_string(memory_budget) +
                                      " bytes): " + hipGetErrorString(e));
    srv->pool = std::make_unique<Pool>((size_t)memory_budget);
    srv->request_cap = request_cap ? (size_t)request_cap : (size_t)memory_budget / 2;
    srv->max_running = max_running;
    for (int32_t i = 0; i < n_streams; ++i) {
        hipStream_t s = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (hipStream_t t : srv->streams) (void)hipStreamDestroy(t);
            (void)hipFree(srv->pool_base);
            return hip_fail(e, "hipStreamCreate");
        }
        srv->streams.push_back(s);
        srv->stream_load.push_back(0);
    }
    mcdp_server *raw = srv.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return MCDP_OK;
}

void mcdp_server_destroy(mcdp_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        srv->stop = true;
        for (Request *r : srv->queued) {  // never started: fail them
            r->state = State::Failed;
            r->rc = MCDP_E_INVALID;
            r->error = "server destroyed before the request started";
        }
        srv->queued.clear();
    }
    srv->cv_done.notify_all();  // waiters on the failed requests return now
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();  // drains running requests
    for (auto &kv : srv->all)
        if (kv.second->h_out) (void)hipHostFree(kv.second->h_out);
    for (hipStream_t s : srv->streams) (void)hipStreamDestroy(s);
    for (double *p : srv->ranges)
        if (p) (void)hipFree(p);
    (void)hipFree(srv->pool_base);
    delete srv;
}

int mcdp_server_add_context(mcdp_server *srv, mcdp_propagator *prop, int32_t *ctx_id) {
    if (!srv || !prop || !ctx_id) return fail(MCDP_E_INVALID, "null argument");
    int rc = prepare_stats(*prop);  // plan and predecessor table resident from now on
    if (rc) return rc;
    // bounds of every event's realized time, for the rows kept whole
    const int32_t entries = max_entries(*prop);
    double *d_ranges = nullptr;
    std::vector<double> lo, hi;
    if (prop->E > 0 && event_ranges(*prop, lo, hi)) {
        std::vector<double> pairs(2 * (size_t)prop->E);
        for (size_t e = 0; e < lo.size(); ++e) {
            pairs[2 * e] = lo[e];
            pairs[2 * e + 1] = hi[e];
        }
        hipError_t e = hipMalloc(&d_ranges, pairs.size() * 8);
        if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (event bounds): ") + hipGetErrorString(e));
        e = hipMemcpy(d_ranges, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d_ranges);
            return hip_fail(e, "upload event bounds");
        }
    }
    std::lock_guard<std::mutex> lock(srv->mu);
    if (srv->contexts.size() >= (size_t)std::numeric_limits<int32_t>::max()) {
        if (d_ranges) (void)hipFree(d_ranges);
        return fail(MCDP_E_INVALID, "too many contexts");
    }
    *ctx_id = (int32_t)srv->contexts.size();
    srv->contexts.push_back(prop);
    srv->ranges.push_back(d_ranges);
    srv->max_entries.push_back(entries);
    return MCDP_OK;
}

int mcdp_server_submit(mcdp_server *srv, int32_t ctx_id, const int32_t *seeds, int64_t n, int32_t n_q,
                       const double *q, int64_t *request_id) {
    if (!srv || !request_id) return fail(MCDP_E_INVALID, "null argument");
    if (n < 1 || !seeds) return fail(MCDP_E_INVALID, "a request needs at least one seed");
    if (n > std::numeric_limits<int32_t>::max() - 256) return fail(MCDP_E_INVALID, "at most 2^31 - 256 seeds per request");
    if (n_q < 0 || n_q > 8 || (n_q > 0 && !q)) return fail(MCDP_E_INVALID, "between 0 and 8 quantiles per request");
    for (int32_t j = 0; j < n_q; ++j)
        if (!(q[j] >= 0.0 && q[j] <= 1.0)) return fail(MCDP_E_INVALID, "quantiles must be in the range [0, 1]");
    auto r = std::make_unique<Request>();
    mcdp_propagator *prop;
    bool exact;
    int32_t kept_limit, entries;
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        if (ctx_id < 0 || (size_t)ctx_id >= srv->contexts.size()) return fail(MCDP_E_INVALID, "unknown context");
        prop = srv->contexts[(size_t)ctx_id];
        exact = srv->exact_chunked;
        kept_limit = srv->kept_limit;
        entries = srv->max_entries[(size_t)ctx_id];
    }
    r->ctx = ctx_id;
    r->E = prop->E;
    r->P = (int64_t)prop->preds.size();
    r->seeds.assign(seeds, seeds + n);
    r->q.assign(q, q + n_q);
    const int cb = cause_bytes_for(entries);
    r->cause_bytes = cb;
    r->bytes = footprint(n, n, r->E, r->P, n_q, false, 0, cb);
    if (r->bytes > srv->request_cap) {
        // too large to hold all samples at once: chunks of samples, the largest that fit the cap
        if (!(prop->max_delay < std::numeric_limits<double>::infinity()))
            return fail(MCDP_E_NOMEM, "request needs " + std::to_string(r->bytes) + " bytes of device memory (cap " +
                                          std::to_string(srv->request_cap) +
                                          "); statistics over sample chunks need a finite max_delay");
        int64_t m, kept;
        chunk_plan(n, r->E, r->P, n_q, srv->request_cap, exact, kept_limit, cb, m, kept);
        if (m < 1)
            return fail(MCDP_E_NOMEM, "request needs " + std::to_string(r->bytes) + " bytes of device memory; even "
                                      "chunks of 128 samples exceed the request cap of " +
                                      std::to_string(srv->request_cap) + " bytes (budget " +
                                      std::to_string(srv->pool->size()) + ")");
        r->chunk = m;
        r->kept_cap = kept;
        r->bytes = footprint(n, r->chunk, r->E, r->P, n_q, true, kept, cb);
    }
    // Admission is strictly in arrival order: a request the pool can never
    // hold would block every later one, so none is queued.
    if (r->bytes > srv->request_cap || r->bytes > srv->pool->size())
        return fail(MCDP_E_NOMEM, "request footprint of " + std::to_string(r->bytes) +
                                      " bytes exceeds the request cap of " + std::to_string(srv->request_cap) +
                                      " bytes");
    // results, then the exact selection's four counters and the kept-row count
    const size_t host_doubles = (size_t)(r->E + r->E * n_q + r->E + r->P);
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&r->h_out), (host_doubles + 8) * 8);
    if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipHostMalloc (results): ") + hipGetErrorString(e));
    r->h_status = reinterpret_cast<uint64_t *>(r->h_out + host_doubles);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        if (srv->stop) {
            (void)hipHostFree(r->h_out);
            return fail(MCDP_E_INVALID, "server is shutting down");
        }
        r->id = srv->next_id++;
        *request_id = r->id;
        srv->queued.push_back(r.get());
        ++srv->n_submitted;
        if (r->chunk > 0) ++srv->n_chunked;
        srv->all.emplace(r->id, std::move(r));
    }
    srv->cv_work.notify_all();
    return MCDP_OK;
}

int mcdp_server_wait(mcdp_server *srv, int64_t request_id, double *mean, double *quantiles, int64_t *cause_counts,
                     double *latency_ms) {
    if (!srv) return fail(MCDP_E_INVALID, "server is null");
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lock(srv->mu);
        auto it = srv->all.find(request_id);
        if (it == srv->all.end()) return fail(MCDP_E_INVALID, "unknown request (or already collected)");
        Request *p = it->second.get();
        srv->cv_done.wait(lock, [&] { return p->state == State::Done || p->state == State::Failed; });
        r = std::move(it->second);
        srv->all.erase(it);
    }
    int rc = r->rc;
    if (r->state == State::Done) {
        const int64_t E = r->E, nq = (int64_t)r->q.size();
        if (mean) std::memcpy(mean, r->h_out, (size_t)E * 8);
        if (quantiles && nq) std::memcpy(quantiles, r->h_out + E, (size_t)(E * nq) * 8);
        if (cause_counts) std::memcpy(cause_counts, r->h_out + E + E * nq, (size_t)(E + r->P) * 8);
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_submit).count();
    } else {
        rc = rc ? rc : MCDP_E_DEVICE;
        (void)fail(rc, r->error.empty() ? "request failed" : r->error);
    }
    (void)hipHostFree(r->h_out);
    return rc;
}

int mcdp_debug_serve_chunk(int64_t n, int64_t n_events, int64_t n_preds, int32_t max_entries, int32_t n_q,
                           int64_t request_cap, int64_t *chunk, int64_t *bytes) {
    if (!chunk || !bytes || n < 1 || n_events < 0 || n_preds < 0 || max_entries < 0 || n_q < 0 || n_q > 8 ||
        request_cap < 0)
        return fail(MCDP_E_INVALID, "invalid argument");
    const size_t cap = (size_t)request_cap;
    const int cb = cause_bytes_for(max_entries);
    const size_t whole = footprint(n, n, n_events, n_preds, n_q, false, 0, cb);
    if (whole <= cap) {
        *chunk = 0;
        *bytes = (int64_t)whole;
        return MCDP_OK;
    }
    int64_t m, kept;
    chunk_plan(n, n_events, n_preds, n_q, cap, true, -1, cb, m, kept);
    if (m < 1) return fail(MCDP_E_NOMEM, "even chunks of 128 samples exceed the request cap");
    *chunk = m;
    *bytes = (int64_t)footprint(n, m, n_events, n_preds, n_q, true, kept, cb);
    return MCDP_OK;
}

int mcdp_server_stats(mcdp_server *srv, int64_t *out, int n_out) {
    if (!srv || !out) return fail(MCDP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(srv->mu);
    const int64_t v[11] = {srv->n_submitted,          srv->n_completed,         srv->n_failed,
                           (int64_t)srv->queued.size(), (int64_t)srv->running.size(), srv->max_running_seen,
                           (int64_t)srv->pool->used(), (int64_t)srv->pool->peak(), srv->n_chunked,
                           srv->n_rounds,               srv->n_kept};
    for (int i = 0; i < n_out && i < 11; ++i) out[i] = v[i];
    return MCDP_OK;
}

int mcdp_server_set_exact_quantiles(mcdp_server *srv, int32_t exact) {
    if (!srv) return fail(MCDP_E_INVALID, "server is null");
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->exact_chunked = exact != 0;
    return MCDP_OK;
}

int mcdp_debug_server_exact_limits(mcdp_server *srv, int32_t c_max, int64_t cand_per_target, int32_t refine_slots) {
    if (!srv) return fail(MCDP_E_INVALID, "server is null");
    if (c_max < 1 || c_max > kExactMaxCollect || cand_per_target < 1 || cand_per_target > kExactCandPerTarget ||
        refine_slots < 1 || refine_slots > kExactRefineSlots)
        return fail(MCDP_E_INVALID, "limits out of range");
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->exact_c_max = c_max;
    srv->exact_cand_per_target = cand_per_target;
    srv->exact_slots = refine_slots;
    return MCDP_OK;
}

int mcdp_debug_server_kept_rows(mcdp_server *srv, int32_t rows) {
    if (!srv) return fail(MCDP_E_INVALID, "server is null");
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->kept_limit = rows < 0 ? -1 : rows;
    return MCDP_OK;
}

}  // extern "C"
