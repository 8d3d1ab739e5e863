/*
 * mcdp.h -- C ABI of the MI355X Monte Carlo DAG propagation library (libmcdp.so).
 *
 * Plain C types only: no C++ or torch types cross this boundary.  Every entry
 * point returns 0 (MCDP_OK) on success or a nonzero mcdp_status; the message of
 * the last failure on the calling thread is available from mcdp_last_error().
 * Bindings turn a nonzero status into the reference's exception type
 * (std::runtime_error -> Python RuntimeError).
 *
 * Each entry point names the reference interface it replaces, as a
 * file:line into WonJayne/mc_dagprop (src/mc_dagprop/monte_carlo/_core.cpp).
 *
 * Layout: per-sample arrays are sample-contiguous ("event-major"):
 *   realized[e * ld + s], cause[e * ld + s]   for event e,   sample s
 *   durations[l * ld + s]                     for activity l, sample s
 * so a wave of 64 samples reads and writes 512-byte runs.  Sample s of the
 * reference's SimResult (_core.cpp:65-69) is column s.
 */
#ifndef MCDP_H
#define MCDP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCDP_ABI_VERSION 1

typedef enum mcdp_status {
    MCDP_OK = 0,
    MCDP_E_INVALID = 1,       /* invalid argument or context                     */
    MCDP_E_RESERVED_TYPE = 2, /* a distribution on activity type -1 (_core.cpp:196-198) */
    MCDP_E_MAX_DELAY = 3,     /* max_delay < 0 (_core.cpp:199-201)                */
    MCDP_E_CYCLE = 4,         /* cycle in precedence list (_core.cpp:262-264)     */
    MCDP_E_DEVICE = 5,        /* no HIP device, or a HIP call failed              */
    MCDP_E_NOMEM = 6          /* host or device allocation failed                 */
} mcdp_status;

typedef struct mcdp_generator mcdp_generator;
typedef struct mcdp_propagator mcdp_propagator;
typedef struct mcdp_batch mcdp_batch;

/* Message of the last failing call on this thread ("" if none). */
const char *mcdp_last_error(void);
/* MCDP_ABI_VERSION of the loaded library. */
int mcdp_abi_version(void);

/* ---- GenericDelayGenerator (_core.cpp:146-159, binding :519-542) -------- */

/* GenericDelayGenerator() (_core.cpp:151). */
int mcdp_generator_create(mcdp_generator **out);
void mcdp_generator_destroy(mcdp_generator *gen);
/* GenericDelayGenerator::set_seed (_core.cpp:153).  Accepted and ignored: the
 * reference never draws from the generator's own engine (:148, :313). */
int mcdp_generator_set_seed(mcdp_generator *gen, int32_t seed);
/* add_constant (_core.cpp:154): extra = factor * minimal_duration. */
int mcdp_generator_add_constant(mcdp_generator *gen, int32_t activity_type, double factor);
/* add_exponential (_core.cpp:155): X ~ Exp(mean lambda_) conditioned on
 * X <= max_scale; extra = X * minimal_duration (:78-90). */
int mcdp_generator_add_exponential(mcdp_generator *gen, int32_t activity_type, double lambda_, double max_scale);
/* add_gamma (_core.cpp:156-158): X ~ Gamma(shape, scale) conditioned on
 * X <= max_scale (pass INFINITY for none); extra = X * minimal_duration. */
int mcdp_generator_add_gamma(mcdp_generator *gen, int32_t activity_type, double shape, double scale,
                             double max_scale);
/* add_empirical_absolute (_core.cpp:529-534, dist :110-126): extra = values[i],
 * P(i) = weights[i] / sum(weights).  n == 1 always draws bin 0. */
int mcdp_generator_add_empirical_absolute(mcdp_generator *gen, int32_t activity_type, const double *values,
                                          const double *weights, int64_t n);
/* add_empirical_relative (_core.cpp:536-541, dist :129-141): extra =
 * factors[i] * minimal_duration. */
int mcdp_generator_add_empirical_relative(mcdp_generator *gen, int32_t activity_type, const double *factors,
                                          const double *weights, int64_t n);

/* ---- MonteCarloPropagator (Simulator, _core.cpp:162-362) ----------------- */

/* Simulator(context, generator) (_core.cpp:193-307): validates, orders and
 * compiles the DAG on the host.  Device upload happens on the first run.
 *   events:      n_events earliest times (EventTimestamp.earliest; latest and
 *                actual are not read by the Monte Carlo engine)
 *   activities:  n_activities (idx, minimal_duration, activity_type); the
 *                (src, dst) keys of the reference's activity map are not read
 *   precedence:  n_entries entries; entry i is (entry_target[i],
 *                [(pred_event[k], pred_activity[k]) for k in
 *                 entry_offsets[i] .. entry_offsets[i+1])]
 * The generator is copied; it may be destroyed afterwards. */
int mcdp_propagator_create(int32_t n_events, const double *earliest, int32_t n_activities,
                           const int32_t *activity_idx, const double *minimal_duration,
                           const int32_t *activity_type, int32_t n_entries, const int32_t *entry_target,
                           const int64_t *entry_offsets, const int32_t *pred_event, const int32_t *pred_activity,
                           double max_delay, const mcdp_generator *gen, mcdp_propagator **out);
void mcdp_propagator_destroy(mcdp_propagator *prop);
/* node_count / activity_count (_core.cpp:309-310). */
int32_t mcdp_propagator_node_count(const mcdp_propagator *prop);
int32_t mcdp_propagator_activity_count(const mcdp_propagator *prop);

/* Simulator::run_many (_core.cpp:355-361) on device buffers, asynchronous on
 * `stream` (a hipStream_t, NULL = default stream).  seeds_dev: int32[n] device;
 * outputs event-major with leading dimension ld >= n.  durations_dev or
 * cause_dev may be NULL to skip that output; realized_dev is required (the
 * kernel reads its own rows back within the launch). */
int mcdp_run_many_device(mcdp_propagator *prop, const int32_t *seeds_dev, int64_t n, int64_t ld,
                         double *realized_dev, double *durations_dev, int32_t *cause_dev, void *stream);

/* Simulator::run_many (_core.cpp:355-361) on host buffers, synchronous.
 * Outputs event-major with ld = n: realized[E*n], durations[L*n], cause[E*n].
 * The device work runs in chunks of samples (about 2 GiB of device buffers,
 * kept by the propagator up to 4 GiB).  Results of 16 MB and more reach the
 * outputs through pinned staging buffers, 64 MB per device, created on the
 * first such call and held until the process ends; calls from several threads
 * take turns on them.  A run of at most 1,024 samples on a plan of 32 steps
 * or more replays a graph of its launches, captured when the same sample count
 * comes twice in a row (the propagator keeps one).  Any output may be NULL to
 * skip it. */
int mcdp_run_many_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, double *realized,
                       double *durations, int32_t *cause);

/* The propagation half of Simulator::run (_core.cpp:316-350) with injected
 * durations (durations_dev[l * ld + s]); device buffers, asynchronous. */
int mcdp_propagate_device(mcdp_propagator *prop, const double *durations_dev, int64_t n, int64_t ld,
                          double *realized_dev, int32_t *cause_dev, void *stream);
/* Same on host buffers (ld = n), synchronous. */
int mcdp_propagate_host(mcdp_propagator *prop, const double *durations, int64_t n, double *realized,
                        int32_t *cause);

/* Plan facts for tests and tuning: out[0..11] = {steps, waves per workgroup,
 * LDS slots, LDS bytes of the slots per workgroup, predecessors read from
 * global memory, stale predecessors, orphan activities, max events per wave
 * and step, events per wave bound, samples per lane, estimated latency of a
 * partial last round as tiles, the same as wide steps (draw-cost units)}.
 * (No counterpart in the reference.) */
int mcdp_propagator_plan_info(const mcdp_propagator *prop, int64_t *out, int n_out);

/* Bounds that hold for every seed: link_lo/link_hi [L] the range of the
 * durations each link can draw, event_lo/event_hi [E] the range of each
 * event's realized time (the propagation recurrence over those ranges).  Any
 * pointer may be NULL.  MCDP_E_INVALID, with the event ranges left unwritten,
 * when an earliest time or max_delay is not finite.  Serving uses them to
 * tell which extremes of a sample chunk later chunks cannot move.  (No
 * counterpart in the reference.) */
int mcdp_propagator_ranges(const mcdp_propagator *prop, double *link_lo, double *link_hi, double *event_lo,
                           double *event_hi);

/* Fixes the schedule shape: `waves` waves per workgroup (1..16), at most
 * `events_per_wave` events per wave and step (4, 8, 12 or 16) and
 * `samples_per_lane` samples per lane (1, or 2 with events_per_wave <= 8).
 * Without it the first run picks a shape for the current device.  Results do
 * not depend on the shape.  (No counterpart in the reference: a tuning knob of
 * the device schedule.) */
int mcdp_propagator_configure(mcdp_propagator *prop, int32_t waves, int32_t events_per_wave,
                              int32_t samples_per_lane);
/* Times every candidate shape on the current device with n_samples samples
 * (seeds 0..n-1, full outputs in temporary device memory) and keeps the
 * fastest; reports it and its time per launch.  (No counterpart in the
 * reference.) */
int mcdp_propagator_autotune(mcdp_propagator *prop, int64_t n_samples, int32_t *waves, int32_t *events_per_wave,
                             int32_t *samples_per_lane, float *ms);
/* A 64-bit fingerprint of what the schedule shape's speed depends on: the
 * precedence structure (events, links, every entry's source, edge and stale
 * flag), each link's distribution and each distribution's kind, and the alias
 * tables' size; not times or distribution parameters (FNV-1a over those
 * values).  (No counterpart in the reference.) */
int mcdp_propagator_fingerprint(const mcdp_propagator *prop, uint64_t *fingerprint);
/* mcdp_propagator_autotune with its choice kept in a text file, one line per
 * choice:  mcdp-shape 1 <fingerprint, 16 hex digits> <device architecture>
 * <compute units> <n_samples rounded up to a power of two> <waves>
 * <events_per_wave> <samples_per_lane> <ms>.  The last line with this
 * propagator's key on the current device is configured without running
 * anything (*from_file = 1); with none (or an invalid or non-resident shape)
 * the autotune runs and appends its line (*from_file = 0), in one write to the
 * file opened for appending.  Lines that do not parse are skipped.  A file
 * that cannot be read counts as empty; one that cannot be written fails the
 * call with MCDP_E_INVALID after the shape is configured.  (No counterpart in
 * the reference.) */
int mcdp_propagator_autotune_cached(mcdp_propagator *prop, int64_t n_samples, const char *path, int32_t *waves,
                                    int32_t *events_per_wave, int32_t *samples_per_lane, float *ms,
                                    int32_t *from_file);

/* ---- Batched run_many over several DAGs (no counterpart in the reference) - */
/* The reference runs one DAG per Simulator; callers run several DAGs from
 * several threads (test/test_simulator.py:201-215).  A batch runs them in one
 * persistent launch: workgroups take (DAG, sample tile) items from one queue,
 * largest DAGs first, so small DAGs and uneven sample counts fill the device.
 * Results are bit-identical to each propagator's own mcdp_run_many_device.
 * A batch holds device state of its own (launch descriptors, a statistics
 * workspace): one batch must not be used from two threads at once. */

/* Makes a batch of n_props propagators (they stay owned by the caller and must
 * outlive the batch).  Every propagator is configured to the batch's schedule
 * shape (waves, events_per_wave, samples_per_lane); all three 0 choose
 * (16, 8, 2).  Reconfiguring one of them afterwards makes runs fail. */
int mcdp_batch_create(mcdp_propagator *const *props, int32_t n_props, int32_t waves, int32_t events_per_wave,
                      int32_t samples_per_lane, mcdp_batch **out);
void mcdp_batch_destroy(mcdp_batch *batch);
/* run_many of every propagator p of the batch on its own device buffers:
 * seeds_dev[p] (n[p] seeds), realized_dev[p] / durations_dev[p] /
 * cause_dev[p] event-major with leading dimension ld[p] >= n[p]
 * (durations_dev / cause_dev may be NULL, or hold NULL entries, to skip those
 * outputs).  Launches on `stream` and returns when the batch is done. */
int mcdp_batch_run_device(mcdp_batch *batch, const int32_t *const *seeds_dev, const int64_t *n, const int64_t *ld,
                          double *const *realized_dev, double *const *durations_dev, int32_t *const *cause_dev,
                          void *stream);
/* mcdp_batch_run_device in two halves, for callers that replay a batch from a
 * HIP graph.  prepare uploads the launch's descriptors (arguments as for
 * run_device; synchronous, waits for the device first, may allocate).  launch
 * enqueues the prepared launch on `stream`: one kernel that clears the work
 * queue's head and the persistent batch kernel, with no allocation, copy or
 * synchronisation, so it may be captured (mcdp_graph_capture_begin/end) and
 * the graph replayed; every replay recomputes every output.  A graph holds the
 * descriptors' device address: the next prepare of the batch changes what its
 * replays compute, and may move the descriptors (then the graph must not be
 * replayed again).  The descriptors also point into each member's device
 * plan, which mcdp_propagator_set_times, _configure and _autotune free: after
 * any of them on a member, launch fails with MCDP_E_INVALID until the batch is
 * prepared again, and a graph captured earlier must not be replayed (capture
 * it again). */
int mcdp_batch_prepare_device(mcdp_batch *batch, const int32_t *const *seeds_dev, const int64_t *n, const int64_t *ld,
                              double *const *realized_dev, double *const *durations_dev, int32_t *const *cause_dev);
int mcdp_batch_launch(mcdp_batch *batch, void *stream);
/* Per-event statistics (as mcdp_event_stats_device defines them) of every DAG
 * p of the batch over its own seeds, without keeping per-sample results: the
 * DAGs run in groups, in batch order, whose seeds, realized and cause rows and
 * statistics fit `memory_budget` bytes of device memory; each group is one
 * batched launch (the other DAGs get no samples) followed by the device
 * statistics of its DAGs.  seeds[p] holds n[p] >= 1 host seeds; results go to
 * host arrays mean[p] (E_p values), quantiles[p] (E_p * n_q; may be NULL when
 * n_q == 0) and cause_counts[p] (off[E_p] + E_p; cause_counts or an entry may
 * be NULL to skip them).  Durations are drawn but not stored.  Synchronous.
 * The device workspace (at most memory_budget bytes) stays with the batch for
 * its next calls until mcdp_batch_destroy.  Fails with MCDP_E_NOMEM if one DAG
 * alone exceeds the budget. */
int mcdp_batch_stats_host(mcdp_batch *batch, const int32_t *const *seeds, const int64_t *n, int32_t n_q,
                          const double *q, int64_t memory_budget, double *const *mean, double *const *quantiles,
                          int64_t *const *cause_counts);

/* ---- Per-event statistics (no counterpart in the reference) -------------- */
/* The reference returns one SimResult per seed and leaves statistics to its
 * callers (demo/monte_carlo.py:45-63 counts realized values per event on the
 * host).  These reduce event-major results on the device instead.  All
 * results are per event e < node_count():
 *   mean[e]                   exactly rounded sum of realized[e, :n], / n
 *   quantiles[e * n_q + j]    numpy.quantile(realized[e, :n], q[j]) (method
 *                             "linear"; NaN if the row holds a NaN), n_q <= 8,
 *                             0 <= q[j] <= 1
 *   cause_counts[off[e] + e]          samples with cause_event -1
 *   cause_counts[off[e] + e + 1 + k]  samples whose cause is the source of the
 *                             k-th predecessor of e (for a source listed twice,
 *                             its first position), off = pred_offsets
 * Any output pointer may be NULL to skip it. */

/* off[0..node_count()]: start of each event's predecessor list (input order of
 * its winning precedence entry); cause_counts holds off[E] + E counters. */
int mcdp_propagator_pred_offsets(const mcdp_propagator *prop, int64_t *off);

/* Statistics of device arrays realized_dev / cause_dev (event-major, leading
 * dimension ld >= n), asynchronous on `stream`; q is a host array; outputs are
 * device arrays. */
int mcdp_event_stats_device(mcdp_propagator *prop, const double *realized_dev, const int32_t *cause_dev, int64_t n,
                            int64_t ld, int32_t n_q, const double *q, double *mean_dev, double *quantiles_dev,
                            int64_t *cause_counts_dev, void *stream);

/* run_many over `seeds` on the device followed by the statistics, without
 * moving per-sample results to the host; outputs are host arrays.
 * Synchronous. */
int mcdp_run_stats_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, int32_t n_q, const double *q,
                        double *mean, double *quantiles, int64_t *cause_counts);

/* ---- Pathwise gradients (no counterpart in the reference) ---------------- */
/* The objective per sample s is
 *     J_s = sum_e weights[e] * max(0, realized[e][s] - earliest[e]),
 * the weighted lateness of every event.  Its pathwise derivative follows the
 * predecessor entry that decided each realized time in the propagation loop
 * (_core.cpp:332-350): an event whose time is its earliest time (cause -1) or
 * its cap earliest + max_delay passes its adjoint to its own earliest time;
 * otherwise to the winning source (or to the source's earliest time when the
 * loop read it before evaluating it) and, through the entry's duration, to
 * the link's minimal duration d: d duration / d d = 1 for links whose type
 * has no distribution or an empirical absolute one, 1 + X for the others
 * (duration = d + X d), evaluated as duration * (1 / d), or by drawing X
 * again when d = 0.  Holds almost everywhere: on a tie the later entry wins,
 * as in the forward loop. */

/* Replaces the minimal duration of every link (dmin[link_count()], indexed
 * by Activity.idx; gaps included) and the earliest time of every event
 * (earliest[node_count()]).  Either may be NULL to keep it.  The schedule is
 * kept; the next launch uploads the new times.  Must not overlap a launch of
 * this propagator. */
int mcdp_propagator_set_times(mcdp_propagator *prop, const double *dmin, const double *earliest);
/* The current times, in the layout of mcdp_propagator_set_times. */
int mcdp_propagator_get_times(const mcdp_propagator *prop, double *dmin, double *earliest);

/* The reverse walk over results of mcdp_run_many_device with the same seeds
 * (event-major, leading dimension ld >= n; durations required).  Writes SUMS
 * over the n samples: objective_dev[0] = sum_s J_s, grad_dmin_dev[l] =
 * sum_s dJ_s/d dmin[l], grad_earliest_dev[e] = sum_s dJ_s/d earliest[e]; with
 * accumulate != 0 it adds them to the outputs' current values.  Device
 * pointers; asynchronous on `stream`.  The propagator's gradient workspace
 * serves one call at a time: calls that overlap on different streams race. */
int mcdp_reverse_device(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev,
                        const double *realized_dev, const double *durations_dev, const int32_t *cause_dev, int64_t n,
                        int64_t ld, int accumulate, double *objective_dev, double *grad_dmin_dev,
                        double *grad_earliest_dev, void *stream);
/* mcdp_reverse_device without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).  Same arguments,
 * results bit-identical to mcdp_reverse_device's.  Its partial rows, unit rows
 * and group sums live in one fixed scratch array of
 * MCDP_NOALLOC_SCRATCH_BYTES per GPU (256 MiB, loaded with the library and
 * held until the process ends), split into as many launches as it needs; the
 * call allocates, frees and synchronises nothing.  Preconditions and limits:
 *   - the propagator's plans must already be resident on the current device
 *     (call mcdp_reverse_device on it once; n = 0 is enough), else
 *     MCDP_E_INVALID;
 *   - a DAG whose reverse walk needs more than the scratch for one launch of
 *     16 tiles (about 1 KiB per partial-row column) fails with MCDP_E_NOMEM;
 *   - every call on a GPU shares that GPU's scratch: calls that overlap on one
 *     GPU (other streams or threads) race and give wrong sums. */
#define MCDP_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int mcdp_reverse_device_noalloc(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev,
                                const double *realized_dev, const double *durations_dev, const int32_t *cause_dev,
                                int64_t n, int64_t ld, int accumulate, double *objective_dev, double *grad_dmin_dev,
                                double *grad_earliest_dev, void *stream);
/* Facts about the reverse walk's plan: out[0..5] = {steps, waves per
 * workgroup, slots, slots in LDS, partial-row columns, samples per lane}. */
int mcdp_propagator_gradient_info(mcdp_propagator *prop, int64_t *out, int n_out);
/* Host convenience: run_many then the reverse walk over chunks of at most
 * `chunk` samples (0: as many as fit a bounded workspace; rounded to whole
 * groups of 16 tiles of the reverse kernel, 1,024 or 2,048 samples), so
 * per-sample results never
 * need to exist for all seeds at once.  Writes MEANS over the n seeds:
 * objective[0], grad_dmin[link_count()], grad_earliest[node_count()].  The
 * result does not depend on the chunk size. */
int mcdp_gradient_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, const double *weights, int64_t chunk,
                       double *objective, double *grad_dmin, double *grad_earliest);

/* ---- Serving (no counterpart in the reference) --------------------------- */
/* A stream of statistics requests over resident contexts.  The reference
 * answers one run_many call at a time on the calling thread
 * (_core.cpp:355-361) and leaves statistics to its callers.  A server owns one
 * device memory pool of `memory_budget` bytes and a scheduler thread.  Each
 * request reserves its footprint (seeds, realized and cause rows, statistics:
 * about 12 * node_count() bytes per sample) from the pool; requests start in
 * arrival order as soon as the pool and one of `max_running` slots allow,
 * while others run, on `n_streams` streams (work of different requests
 * overlaps on the device).  A request's results equal mcdp_run_stats_host on
 * the same seeds (for chunked requests, see request_cap).  All entry points are
 * thread-safe. */
typedef struct mcdp_server mcdp_server;
/* request_cap (0: memory_budget / 2) bounds one request's share of the pool:
 * a request needing more runs in chunks of samples that fit it (needs a finite
 * max_delay).  Its mean and cause counts come from running sums over the
 * chunks.  Its quantiles: an order statistic on the mass of an extreme is
 * that extreme; the events whose first chunk puts an order statistic
 * elsewhere (as far as a chunk tells, with the event's bounds from
 * mcdp_propagator_ranges saying which extremes later chunks cannot move)
 * keep their whole rows across the chunks, and their quantiles come from
 * those rows; any other order statistic takes further passes over the same
 * chunks (results are a function of the seeds) that gather, or narrow down
 * in 256 sub-bins, the values around it that a 4,096-bin histogram over
 * [earliest, earliest + max_delay] placed.  Every statistic equals
 * mcdp_run_stats_host's bit for bit.  Usually no further pass is needed;
 * each pass shrinks a region that holds too many values 256-fold. */
int mcdp_server_create(int64_t memory_budget, int64_t request_cap, int32_t max_running, int32_t n_streams,
                       mcdp_server **out);
/* Waits for running requests; fails queued ones; frees the pool. */
void mcdp_server_destroy(mcdp_server *srv);
/* Makes a context resident (plan and predecessor table uploaded now) and
 * returns its id.  The propagator must outlive the server. */
int mcdp_server_add_context(mcdp_server *srv, mcdp_propagator *prop, int32_t *ctx_id);
/* Queues run_many over seeds[0..n) of context ctx_id and the statistics of
 * mcdp_run_stats_host (n_q <= 8 quantiles); returns at once.  Fails with
 * MCDP_E_NOMEM if even chunks of 128 samples cannot fit the request cap. */
int mcdp_server_submit(mcdp_server *srv, int32_t ctx_id, const int32_t *seeds, int64_t n, int32_t n_q,
                       const double *q, int64_t *request_id);
/* Blocks until the request is done and writes its results (layouts of
 * mcdp_run_stats_host; any may be NULL) and its latency from submission to
 * completion; the request is then forgotten. */
int mcdp_server_wait(mcdp_server *srv, int64_t request_id, double *mean, double *quantiles, int64_t *cause_counts,
                     double *latency_ms);
/* out[0..10] = {submitted, completed, failed, queued, running, most running at
 * once, pool bytes in use, peak pool bytes in use, requests run in chunks,
 * further passes over chunks for exact quantiles, rows kept whole}. */
int mcdp_server_stats(mcdp_server *srv, int64_t *out, int n_out);
/* exact != 0 (the default): chunked requests get exact quantiles (further
 * passes over their chunks).  exact == 0: no further passes; each quantile is
 * then within max_delay / 4096 of the exact one (the order statistics placed
 * inside their histogram bin), for lower latency. */
int mcdp_server_set_exact_quantiles(mcdp_server *srv, int32_t exact);

/* ---- Device plumbing for callers without their own HIP allocator -------- */
/* (no counterpart in the reference, which has no device) */
int mcdp_device_count(int *count);
int mcdp_malloc(void **ptr, size_t bytes);
int mcdp_free(void *ptr);
int mcdp_memcpy_htod(void *dst_dev, const void *src_host, size_t bytes);
int mcdp_memcpy_dtoh(void *dst_host, const void *src_dev, size_t bytes);
int mcdp_memset(void *dst_dev, int value, size_t bytes);
/* Copies `height` rows of `width` bytes from a device array with row pitch
 * `src_pitch` to a host array with row pitch `dst_pitch` (e.g. a block of
 * sample columns out of an event-major output). */
int mcdp_memcpy2d_dtoh(void *dst_host, size_t dst_pitch, const void *src_dev, size_t src_pitch, size_t width,
                       size_t height);
int mcdp_stream_create(void **stream);
int mcdp_stream_destroy(void *stream);
int mcdp_stream_synchronize(void *stream);
int mcdp_event_create(void **event);
int mcdp_event_destroy(void *event);
int mcdp_event_record(void *event, void *stream);
int mcdp_event_elapsed_ms(void *start, void *stop, float *ms);
/* Stream capture into a HIP graph: begin on `stream` (global capture mode),
 * end returns an instantiated graph; launch replays it on a stream; destroy
 * frees it. */
int mcdp_graph_capture_begin(void *stream);
int mcdp_graph_capture_end(void *stream, void **graph_exec);
int mcdp_graph_launch(void *graph_exec, void *stream);
int mcdp_graph_destroy(void *graph_exec);

/* ---- Test hooks (no counterpart in the reference) ------------------------ */
/* Runs the device Philox-4x32-10 (variant 4) or Philox-2x32-10 (variant 2) on
 * n host-supplied counters/keys and copies the words back to the host.
 * variant 4: ctr[4n], key[2n], out[4n]; variant 2: ctr[2n], key[n], out[2n]. */
int mcdp_debug_philox(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);
/* The samplers' device elementary functions on n host inputs, results copied
 * back: fn 0 = natural log (x > 0), 1 = cos(2 pi x) (x in [0, 1]), 2 = e^x
 * (x <= 0). */
int mcdp_debug_math(int fn, int64_t n, const double *x, double *out);
/* Plain device copy of n doubles (one read and one write per element): the
 * calibration kernel for HBM byte counters. */
int mcdp_debug_copy(double *dst_dev, const double *src_dev, int64_t n, void *stream);
/* Caps the tiles one reverse-kernel launch of this propagator covers (0: the
 * default, bounded by a 1 GiB workspace; rounded down to whole groups of 16
 * tiles, at least one), so tests can run a mcdp_reverse_device call as several
 * launches. */
int mcdp_debug_reverse_tiles(mcdp_propagator *prop, int64_t max_tiles);
/* Route of a launch's last partial round (the samples past its last full
 * round of tiles): 0 = chosen from the schedule's latency estimates (the
 * default), 1 = tiles of one sample per lane, 2 = wide steps (one launch per
 * step, every event a wave per group of 64 samples).  Results do not depend on
 * the route; tests use it to run both. */
int mcdp_debug_tail_mode(mcdp_propagator *prop, int32_t mode);
/* The serving footprint rule, host only: for a request of n samples over a
 * context of n_events events and n_preds predecessor entries (at most
 * max_entries for one event: up to 254, chunks store one-byte winner
 * positions instead of four-byte cause events) with n_q quantiles, writes
 * chunk = 0 and the request's device bytes when it fits request_cap whole,
 * else the samples per chunk and the chunked request's bytes (<= request_cap);
 * fails with MCDP_E_NOMEM when not even chunks of 128 samples fit, as
 * mcdp_server_submit does. */
int mcdp_debug_serve_chunk(int64_t n, int64_t n_events, int64_t n_preds, int32_t max_entries, int32_t n_q,
                           int64_t request_cap, int64_t *chunk, int64_t *bytes);
/* Lowers the exact selection's limits for later requests: values gathered per
 * region (1..4096), candidate space per target (1..256 values) and refine
 * slots per request (1..1024), so tests reach the refining passes and the
 * targets that wait for a slot. */
int mcdp_debug_server_exact_limits(mcdp_server *srv, int32_t c_max, int64_t cand_per_target, int32_t refine_slots);
/* Rows kept whole per chunked request, for later requests: rows < 0 restores
 * the default (one per 128 events, at least 16, within a quarter of the
 * request cap), 0 keeps none (every order statistic off the extremes' masses
 * then takes further passes), rows > 0 at most that many. */
int mcdp_debug_server_kept_rows(mcdp_server *srv, int32_t rows);

#ifdef __cplusplus
}
#endif
#endif /* MCDP_H */
