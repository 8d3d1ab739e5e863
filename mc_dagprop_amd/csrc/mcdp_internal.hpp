// Host-side internals of libmcdp shared by its translation units: the
// propagator object behind the C handle, its per-device copies, and the error
// helpers that set mcdp_last_error().
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_grad.hpp"
#include "mcdp_kernels.hpp"
#include "mcdp_plan.hpp"
#include "mcdp_stats.hpp"

namespace mcdp {

// Sets the calling thread's last error message; returns `code`.
int fail(int code, const std::string &msg);
// fail(MCDP_E_DEVICE, "<what>: <HIP error string>").
int hip_fail(hipError_t e, const char *what);

#define HIP_TRY(expr, what)                                        \
    do {                                                           \
        hipError_t _e = (expr);                                    \
        if (_e != hipSuccess) return ::mcdp::hip_fail(_e, what);   \
    } while (0)

// A compiled distribution (kind, flags, parameters): records of edges of its
// activity type carry kind, flags and p0 inline and its table index.
using DistRec = DistParams;
constexpr uint32_t kNoDistIndex = 0xFFFFFFFFu;  // the link's type has no distribution

// The plan (and, once statistics are asked for, the predecessor table) on one
// device.
struct DeviceCopy {
    int device = -1;
    void *blob = nullptr;
    WalkParams base{};
    void *topo = nullptr;                // statistics: predecessor table, uploaded on first use
    const int32_t *pred_off = nullptr;   // [E + 1]
    const int32_t *pred_key = nullptr;   // [P] per event: its distinct sources, ascending
    const int32_t *pred_pos = nullptr;   // [P] first position of pred_key in the event's list
    const int32_t *pred_nkey = nullptr;  // [E] distinct sources per event
    const int32_t *pred_first = nullptr; // [P] per entry: the first position of its source in the event's list
    void *rev = nullptr;                 // gradients: reverse plan, uploaded on first use
    RevParams rev_base{};
    const int32_t *rev_map_off = nullptr;  // [L + E + 2] columns of each output
    const int32_t *rev_map_col = nullptr;
    const Record *rev_unit_links = nullptr;
    void *rev_work = nullptr;            // partial rows and global slots
    size_t rev_work_bytes = 0;
    // A second stream for the wide steps of a last partial round, which run
    // beside the tiles of the full rounds (created on first use), and the
    // events that fork it from and join it to the caller's stream.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// The reverse walk's plan (mcdp_grad.cpp): the forward levels in reverse
// order, split into steps for `waves` waves of at most kRevEventsPerWave
// events, with the records, slots and partial-row columns of each event.
struct RevPlan {
    int32_t waves = 1, n_steps = 0, slots = 0, n_cols = 0, obj_col = 0;
    int32_t spl = 1;  // samples per lane of the reverse kernel
    std::vector<int32_t> step_off, ev_off, ev_list, ev_rec, ev_succ, ev_slot;
    std::vector<int32_t> succ;     // (target slot, position) pairs
    std::vector<int32_t> col_off;  // first partial-row column of each (step, wave)
    std::vector<Record> recs;
    std::vector<int32_t> map_off, map_col;  // outputs: L links, E events, the objective
    std::vector<Record> unit_links;         // forward records of the links drawn again (kMulDraw)
};

// Device buffers and a stream of the host-buffer entry points
// (mcdp_run_many_host, mcdp_propagate_host) on one device: kept for the next
// call, so a call allocates only when it needs more than any call before it,
// and waits on its own stream only.  They hold per-sample rows, no plan, so
// new times or a new shape leave them valid.
// What a short run's launches hold fixed: the workspace, the sample count and
// leading dimension (which place every buffer in it), the mode, the plan's
// generation and the route hook.  Equal keys launch the same kernels with the
// same arguments.
struct ShortRunKey {
    const void *buf = nullptr;
    int64_t n = -1, ld = -1;
    bool injected = false;
    uint64_t generation = 0;
    int32_t tail_mode = 0;
    bool operator==(const ShortRunKey &o) const {
        return buf == o.buf && n == o.n && ld == o.ld && injected == o.injected && generation == o.generation &&
               tail_mode == o.tail_mode;
    }
};

struct HostWork {
    int device = -1;
    void *buf = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    // Short runs: the launches of a key seen twice in a row are captured into
    // a graph, replayed while the key stays (one graph launch instead of a
    // launch per step).
    ShortRunKey seen, captured;
    hipGraphExec_t graph = nullptr;
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        seen = captured = ShortRunKey{};
    }
};

// One predecessor entry as the reference evaluates it.
struct PredEntry {
    int32_t src;
    int32_t edge;
    bool stale;  // read before src is evaluated in the reference's order
};

}  // namespace mcdp

struct mcdp_propagator {
    int32_t E = 0, L = 0;
    double max_delay = 0.0;
    // topology (schedule-independent)
    std::vector<double> earliest;              // [E]
    std::vector<std::vector<int32_t>> levels;  // events by level, in Kahn order
    std::vector<int64_t> pred_off;             // [E+1] into preds (winning entry, input order)
    std::vector<mcdp::PredEntry> preds;
    std::vector<double> dmin;                  // [L]
    std::vector<uint32_t> ldist;               // [L] distribution index or kNoDistIndex
    std::vector<mcdp::DistRec> dist;
    std::vector<mcdp::AliasRec> alias;
    int64_t n_stale = 0;
    // schedule (for `waves` waves per workgroup, at most `k` events per wave and step)
    std::vector<int32_t> step_off, ev_off, slot_list;
    std::vector<int32_t> ev_rec;  // header record of each event, in slot_list order
    std::vector<mcdp::Record> recs, orph;
    // Estimated latency of the last partial round, in draw-cost units: as
    // tiles (sum over steps of the busiest wave's events) and as wide steps
    // (sum over steps of the costliest event plus the launch gap).
    int64_t tile_cost = 0, wide_cost = 0;
    int32_t tail_mode = 0;  // test hook: 0 by the estimates, 1 tiles, 2 wide steps
    int32_t n_steps = 0, waves = 1, slots = 0, max_per_wave = 0, k = mcdp::kMaxEventsPerWave;
    int32_t spl = 1;  // samples per lane of the walk kernel (schedule-independent)
    int64_t n_global = 0;
    bool shape_fixed = false;  // set by configure / autotune / the first launch
    std::shared_ptr<mcdp::RevPlan> rev;  // built on the first gradient call
    int64_t rev_max_tiles = 0;           // test hook: tiles per reverse launch (0: default)
    // device copies, one per device used
    std::mutex mu;
    std::vector<mcdp::DeviceCopy> dev;
    // Bumped whenever the device copies are dropped (new times, a new shape):
    // a batch launch prepared against older copies would read freed plans.
    std::atomic<uint64_t> generation{0};

    // the host-buffer entry points' buffers and streams, one per device; a
    // call holds host_mu throughout (taken before mu)
    std::mutex host_mu;
    std::vector<mcdp::HostWork> host_work;

    void drop_device_copies() {
        generation.fetch_add(1);
        for (auto &d : dev) {
            int cur = -1;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(d.device);
            (void)hipDeviceSynchronize();  // launches in flight on any stream may still read the plan
            if (d.blob) (void)hipFree(d.blob);
            if (d.topo) (void)hipFree(d.topo);
            if (d.rev) (void)hipFree(d.rev);
            if (d.rev_work) (void)hipFree(d.rev_work);
            if (d.side) (void)hipStreamDestroy(d.side);
            if (d.fork) (void)hipEventDestroy(d.fork);
            if (d.join) (void)hipEventDestroy(d.join);
            if (cur >= 0) (void)hipSetDevice(cur);
        }
        dev.clear();
    }
    ~mcdp_propagator() {
        drop_device_copies();
        for (auto &h : host_work) {
            int cur = -1;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(h.device);
            if (h.stream) (void)hipStreamSynchronize(h.stream);
            h.drop_graph();
            if (h.buf) (void)hipFree(h.buf);
            if (h.stream) (void)hipStreamDestroy(h.stream);
            if (cur >= 0) (void)hipSetDevice(cur);
        }
    }
};

namespace mcdp {
// The propagator's plan on the current device (uploaded on first use).  The
// caller holds P.mu.
int ensure_device(mcdp_propagator &P, DeviceCopy *&out);
// Walk-kernel variant (fast or general paths) and its LDS bytes for P's plan.
WalkVariant plan_variant(const mcdp_propagator &P, bool injected);
size_t plan_lds_bytes(const mcdp_propagator &P, WalkVariant v);
// The host-buffer entry points' device workspace of at least `need` bytes and
// stream on the current device (made or grown on demand).  The caller holds
// P.host_mu until its work on the stream is done, and ends by waiting on it.
int host_workspace(mcdp_propagator &P, size_t need, void *&buf, hipStream_t &stream);
// Frees the current device's workspace if it is larger than
// kHostWorkKeepBytes (a large call does not pin its buffers for good); the
// caller holds P.host_mu and its stream is idle.
constexpr size_t kHostWorkKeepBytes = (size_t)4 << 30;
void host_workspace_trim(mcdp_propagator &P);
// mcdp_reverse_device with its workspace given: `work` of `work_bytes` bytes,
// split into as many launches as it needs, and nothing allocated (the
// propagator's plans must be resident on the current device already); or,
// with work == nullptr, the propagator's own workspace, grown on demand.
int reverse_run(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev, const double *realized_dev,
                const double *durations_dev, const int32_t *cause_dev, int64_t n, int64_t ld, int accumulate,
                double *objective_dev, double *grad_dmin_dev, double *grad_earliest_dev, void *work, size_t work_bytes,
                hipStream_t s);
// mcdp_propagator_configure.
int configure_shape(mcdp_propagator *prop, int32_t waves, int32_t k, int32_t spl);
// Uploads P's plan and the statistics' predecessor table to the current
// device (what a resident context needs before serving requests).
int prepare_stats(mcdp_propagator &P);
// Statistics of a run in sample chunks (serving): adds one chunk's rows
// (realized[e * ld + s], cause likewise, or cause_pos when not null: winner
// positions; s < m) to the running statistics and cause counts (first: start
// them), then the means and quantiles of all n samples; asynchronous on s.
int chunk_stats_accumulate(mcdp_propagator &P, const double *realized, const int32_t *cause, const uint8_t *cause_pos,
                           int32_t m, int32_t ld, bool first, ChunkStat *st, uint32_t *hist, int64_t *counts,
                           hipStream_t s);
int chunk_stats_finish(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                       const uint32_t *hist, double *mean, double *quantiles, hipStream_t s);
// Exact quantiles of a run in sample chunks (mcdp_stats.hpp, ExactWork): the
// targets from the running statistics (w.n_ranks set from n and q), one chunk
// of a further pass, the end of a pass, and the quantiles once every target
// is resolved; asynchronous on s.  The number of distinct ranks n and q need.
int exact_rank_count(int64_t n, int32_t n_q, const double *q, int32_t *n_ranks);
int exact_stats_init(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                     const uint32_t *hist, ExactWork &w, const WatchWork *watch, hipStream_t s);
int exact_stats_pass(mcdp_propagator &P, const double *realized, int32_t m, int32_t ld, const ExactWork &w,
                     hipStream_t s);
int exact_stats_round_end(mcdp_propagator &P, const ExactWork &w, hipStream_t s);
int exact_stats_finish(mcdp_propagator &P, int64_t n, int32_t n_q, const double *q, const ChunkStat *st,
                       const ExactWork &w, double *quantiles, hipStream_t s);
// mcdp_event_stats_device with winner positions (run_many_positions) in place
// of cause events; the same results.
int event_stats_positions(mcdp_propagator &P, const double *realized, const uint8_t *cause_pos, int64_t n, int64_t ld,
                          int32_t n_q, const double *q, double *mean, double *quantiles, int64_t *counts,
                          hipStream_t s);
// Statistics of several DAGs in one launch per kernel (RowSet, mcdp_stats.hpp):
// one DAG's set (rows realized / cause with ld = n, its global rows from
// row0, the outputs of mcdp_event_stats_device; counts null: none), then the
// launch over sets already in device memory; asynchronous on s.
int row_set(mcdp_propagator &P, const double *realized, const int32_t *cause, const uint8_t *cause_pos, int64_t n,
            int32_t n_q, const double *q, double *mean, double *quantiles, int64_t *counts, int64_t row0, RowSet &rs);
int set_stats(const RowSet *sets_dev, int32_t n_sets, int64_t rows, bool quantiles, bool counts, hipStream_t s);
// Rows kept whole across a request's chunks (mcdp_stats.hpp, WatchWork): c[0..n)
// = 0; the rows chosen after the first chunk (ranges: [E][2] event bounds);
// one chunk's samples c0 .. c0 + m appended; and the quantiles of the first
// `rows` kept rows (ld = n) written over theirs in quantiles[E][n_q].
// Asynchronous on s.
int counters_clear(uint64_t *c, int n, hipStream_t s);
int watch_select(mcdp_propagator &P, int64_t n, int32_t m, int32_t n_q, const double *q, const ChunkStat *st,
                 const double *ranges, const WatchWork &w, hipStream_t s);
int watch_copy(const double *realized, int32_t m, int32_t ld, int64_t c0, const WatchWork &w, hipStream_t s);
int watch_finish(int64_t n, int32_t n_q, const double *q, const WatchWork &w, int32_t rows, double *quantiles,
                 hipStream_t s);
// The forward plan's record of link l (kind, flags, distribution index,
// minimal duration, p0); an edge outside the link table is kind kNone.
Record link_record(const mcdp_propagator &P, int32_t l);
// run_many with each event's winning entry position (one byte per sample,
// 0xFF: none) in place of its cause event, for events of at most
// kMaxPositionEntries entries (max_entries: the most any event has);
// asynchronous on s.
constexpr int32_t kMaxPositionEntries = 254;
int32_t max_entries(const mcdp_propagator &P);
int run_many_positions(mcdp_propagator &P, const int32_t *seeds_dev, int64_t n, int64_t ld, double *realized,
                       uint8_t *cause_pos, hipStream_t s);
// Bounds that hold for every seed: the durations link l can draw, and every
// event's realized time (false when an earliest time or max_delay is not
// finite).
void link_duration_range(const mcdp_propagator &P, int32_t l, double &lo, double &hi);
bool event_ranges(const mcdp_propagator &P, std::vector<double> &ev_lo, std::vector<double> &ev_hi);
}  // namespace mcdp
