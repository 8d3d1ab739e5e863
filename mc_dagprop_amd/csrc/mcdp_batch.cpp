// Batched run_many over several compiled DAGs in one persistent launch
// (C ABI in include/mcdp.h).  The reference runs one DAG per Simulator and
// leaves parallelism across DAGs to threads (test/test_simulator.py:201-215);
// here workgroups take (DAG, sample tile) items from one queue, largest DAGs
// first, so small DAGs and uneven sample counts still fill the device.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"

using namespace mcdp;

struct mcdp_batch {
    std::vector<mcdp_propagator *> props;
    int32_t waves = 16, k = 8, spl = 2;
    void *dev = nullptr;  // WalkParams[plans], item_dag[], item_tile[], counter
    size_t dev_bytes = 0;
    int device = -1;
    // The prepared launch (mcdp_batch_prepare_device): descriptors in `dev`,
    // and the generation of each member's device copies they point into.
    bool prepared = false;
    std::vector<uint64_t> gens;
    WalkBatchParams bp{};
    WalkVariant v{};
    size_t lds = 0;
    int grid = 0;
    // mcdp_batch_stats_host's workspace, kept for the next call (allocating
    // hundreds of GB costs more than the launches that use them)
    void *stats_work = nullptr;
    size_t stats_bytes = 0;
    int stats_device = -1;
    ~mcdp_batch() {
        if (dev || stats_work) (void)hipDeviceSynchronize();
        if (dev) (void)hipFree(dev);
        if (stats_work) (void)hipFree(stats_work);
    }
};

namespace {
size_t align_to(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

extern "C" {

int mcdp_batch_create(mcdp_propagator *const *props, int32_t n_props, int32_t waves, int32_t events_per_wave,
                      int32_t samples_per_lane, mcdp_batch **out) {
    if (!out) return fail(MCDP_E_INVALID, "out is null");
    *out = nullptr;
    if (n_props < 1 || !props) return fail(MCDP_E_INVALID, "a batch needs at least one propagator");
    for (int32_t i = 0; i < n_props; ++i)
        if (!props[i]) return fail(MCDP_E_INVALID, "propagator is null");
    if (waves == 0 && events_per_wave == 0 && samples_per_lane == 0) {
        waves = 16;
        events_per_wave = 8;
        samples_per_lane = 2;
    }
    if (samples_per_lane != 1 && samples_per_lane != 2)
        return fail(MCDP_E_INVALID, "samples_per_lane must be 1 or 2");
    for (int32_t i = 0; i < n_props; ++i) {  // one schedule shape for every plan of the launch
        int rc = configure_shape(props[i], waves, events_per_wave, samples_per_lane);
        if (rc) return rc;
    }
    auto *b = new (std::nothrow) mcdp_batch;
    if (!b) return fail(MCDP_E_NOMEM, "out of host memory");
    b->props.assign(props, props + n_props);
    b->waves = waves;
    b->k = events_per_wave;
    b->spl = samples_per_lane;
    *out = b;
    return MCDP_OK;
}

void mcdp_batch_destroy(mcdp_batch *b) { delete b; }

}  // extern "C"

namespace {

// mcdp_batch_prepare_device; cause_pos_dev (instead of cause_dev, every plan
// of at most kMaxPositionEntries entries per event): the winner positions.
int prepare(mcdp_batch *b, const int32_t *const *seeds_dev, const int64_t *n, const int64_t *ld,
            double *const *realized_dev, double *const *durations_dev, int32_t *const *cause_dev,
            uint8_t *const *cause_pos_dev) {
    if (!b) return fail(MCDP_E_INVALID, "batch is null");
    if (cause_pos_dev && cause_dev) return fail(MCDP_E_INVALID, "internal error: positions replace causes");
    b->prepared = false;
    if (!seeds_dev || !n || !ld || !realized_dev) return fail(MCDP_E_INVALID, "null argument");
    const size_t D = b->props.size();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");

    // Per-plan launch parameters; the fast variant only if every plan allows it.
    std::vector<WalkParams> params(D);
    std::vector<int64_t> tiles(D), cost(D);
    std::vector<uint64_t> gens(D);
    bool all_fast = true;
    size_t lds = 0;
    const int64_t tile = (int64_t)kLanes * b->spl;
    for (size_t d = 0; d < D; ++d) {
        mcdp_propagator &P = *b->props[d];
        if (n[d] < 0 || (n[d] > 0 && ld[d] < n[d])) return fail(MCDP_E_INVALID, "need 0 <= n <= ld for every plan");
        if (n[d] > std::numeric_limits<int32_t>::max() - 256 || ld[d] > std::numeric_limits<int32_t>::max())
            return fail(MCDP_E_INVALID, "n and ld must be below 2^31 - 256 per plan");
        if (n[d] > 0 && (!seeds_dev[d] || (P.E > 0 && !realized_dev[d])))
            return fail(MCDP_E_INVALID, "seeds and realized are required for every plan with samples");
        std::lock_guard<std::mutex> lock(P.mu);
        if (P.waves != b->waves || P.k != b->k || P.spl != b->spl)
            return fail(MCDP_E_INVALID, "a propagator of the batch was reconfigured after the batch was made");
        DeviceCopy *dc = nullptr;
        int rc = ensure_device(P, dc);
        if (rc) return rc;
        gens[d] = P.generation.load();
        WalkParams p = dc->base;
        p.seeds = seeds_dev[d];
        p.dur_in = nullptr;
        p.n = (int32_t)n[d];
        p.ld = (int32_t)std::max<int64_t>(ld[d], 1);
        p.realized = realized_dev[d];
        p.durations = durations_dev ? durations_dev[d] : nullptr;
        p.cause = cause_dev ? cause_dev[d] : nullptr;
        p.cause_pos = cause_pos_dev ? cause_pos_dev[d] : nullptr;
        if (p.cause_pos && max_entries(P) > kMaxPositionEntries)
            return fail(MCDP_E_INVALID, "internal error: winner positions need events of at most 254 entries");
        const bool has_work = n[d] > 0 && (P.E > 0 || (!P.orph.empty() && p.durations));
        tiles[d] = has_work ? (n[d] + tile - 1) / tile : 0;
        p.n_tiles = (int32_t)tiles[d];
        params[d] = p;
        const WalkVariant v = plan_variant(P, false);
        all_fast = all_fast && v.paths == kPathsFast;
        cost[d] = (int64_t)P.recs.size() + (int64_t)P.orph.size();
    }
    WalkVariant v{false, all_fast ? kPathsFast : kPathsGeneral, b->k, b->spl};
    v.pos = cause_pos_dev != nullptr;
    for (size_t d = 0; d < D; ++d) lds = std::max(lds, plan_lds_bytes(*b->props[d], v));

    // Work items: every tile of every plan, plans with the most records first.
    std::vector<size_t> order(D);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<int32_t> item_dag, item_tile;
    for (size_t d : order)
        for (int64_t t = 0; t < tiles[d]; ++t) {
            item_dag.push_back((int32_t)d);
            item_tile.push_back((int32_t)t);
        }
    if (item_dag.size() > (size_t)std::numeric_limits<int32_t>::max() / 2)
        return fail(MCDP_E_INVALID, "too many tiles in one batch");
    const int32_t n_items = (int32_t)item_dag.size();

    const size_t o_par = 0, o_dag = align_to(sizeof(WalkParams) * D, 256);
    const size_t o_tile = o_dag + align_to(sizeof(int32_t) * item_dag.size(), 256);
    const size_t o_cnt = o_tile + align_to(sizeof(int32_t) * item_tile.size(), 256);
    const size_t bytes = o_cnt + 256;
    if (!b->dev || b->dev_bytes < bytes || b->device != device) {
        if (b->dev) {
            (void)hipDeviceSynchronize();
            (void)hipFree(b->dev);
            b->dev = nullptr;
        }
        hipError_t e = hipMalloc(&b->dev, bytes);
        if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (batch): ") + hipGetErrorString(e));
        b->dev_bytes = bytes;
        b->device = device;
    }
    auto *base = static_cast<unsigned char *>(b->dev);
    // Launches in flight may still read the descriptors being replaced.
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_TRY(hipMemcpy(base + o_par, params.data(), sizeof(WalkParams) * D, hipMemcpyHostToDevice), "copy batch");
    if (n_items > 0) {
        HIP_TRY(hipMemcpy(base + o_dag, item_dag.data(), sizeof(int32_t) * item_dag.size(), hipMemcpyHostToDevice),
                "copy batch");
        HIP_TRY(hipMemcpy(base + o_tile, item_tile.data(), sizeof(int32_t) * item_tile.size(), hipMemcpyHostToDevice),
                "copy batch");
    }
    WalkBatchParams bp;
    bp.params = reinterpret_cast<const WalkParams *>(base + o_par);
    bp.item_dag = reinterpret_cast<const int32_t *>(base + o_dag);
    bp.item_tile = reinterpret_cast<const int32_t *>(base + o_tile);
    bp.n_items = n_items;
    bp.counter = reinterpret_cast<int32_t *>(base + o_cnt);
    int cus = 0, bpc = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    HIP_TRY(walk_batch_occupancy(v, b->waves, lds, &bpc), "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (bpc < 1) return fail(MCDP_E_DEVICE, "batch kernel cannot be resident with " + std::to_string(lds) + " bytes of LDS");
    b->bp = bp;
    b->v = v;
    b->lds = lds;
    b->grid = (int)std::min<int64_t>(n_items, (int64_t)cus * bpc);
    b->gens = std::move(gens);
    b->prepared = true;
    return MCDP_OK;
}

}  // namespace

extern "C" {

int mcdp_batch_prepare_device(mcdp_batch *b, const int32_t *const *seeds_dev, const int64_t *n, const int64_t *ld,
                              double *const *realized_dev, double *const *durations_dev, int32_t *const *cause_dev) {
    return prepare(b, seeds_dev, n, ld, realized_dev, durations_dev, cause_dev, nullptr);
}

int mcdp_batch_launch(mcdp_batch *b, void *stream) {
    if (!b) return fail(MCDP_E_INVALID, "batch is null");
    if (!b->prepared) return fail(MCDP_E_INVALID, "no prepared launch: call mcdp_batch_prepare_device first");
    // set_times / configure / autotune of a member free the plan the prepared
    // descriptors point into: such a launch (or a graph captured from it) is
    // refused until the batch is prepared again
    for (size_t d = 0; d < b->props.size(); ++d)
        if (b->props[d]->generation.load() != b->gens[d]) {
            b->prepared = false;
            return fail(MCDP_E_INVALID, "a propagator of the batch changed its times or shape since "
                                        "mcdp_batch_prepare_device: prepare the batch again (and re-capture graphs)");
        }
    if (b->bp.n_items == 0) return MCDP_OK;
    HIP_TRY(launch_walk_batch(b->bp, b->v, b->waves, b->grid, b->lds, static_cast<hipStream_t>(stream)),
            "batch kernel launch");
    return MCDP_OK;
}

int mcdp_batch_run_device(mcdp_batch *b, const int32_t *const *seeds_dev, const int64_t *n, const int64_t *ld,
                          double *const *realized_dev, double *const *durations_dev, int32_t *const *cause_dev,
                          void *stream) {
    int rc = mcdp_batch_prepare_device(b, seeds_dev, n, ld, realized_dev, durations_dev, cause_dev);
    if (rc) return rc;
    rc = mcdp_batch_launch(b, stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return MCDP_OK;
}

int mcdp_batch_stats_host(mcdp_batch *b, const int32_t *const *seeds, const int64_t *n, int32_t n_q, const double *q,
                          int64_t memory_budget, double *const *mean, double *const *quantiles,
                          int64_t *const *cause_counts) {
    if (!b) return fail(MCDP_E_INVALID, "batch is null");
    if (!seeds || !n || !mean) return fail(MCDP_E_INVALID, "null argument");
    if (n_q < 0 || (n_q > 0 && (!q || !quantiles))) return fail(MCDP_E_INVALID, "quantiles need q and outputs");
    if (memory_budget <= 0) return fail(MCDP_E_INVALID, "memory_budget must be positive");
    const size_t D = b->props.size();
    const size_t a = 256;
    // The walks store one-byte winner positions instead of cause events when
    // no event of any DAG has more than 254 entries (the same counts).
    bool pos = true;
    for (size_t d = 0; d < D; ++d) pos = pos && max_entries(*b->props[d]) <= kMaxPositionEntries;
    const size_t cb = pos ? 1 : 4;
    // Device bytes of one DAG's part of a group: seeds, realized and cause
    // rows (ld = n), and its statistics.
    std::vector<size_t> bytes(D), o_seed(D), o_real(D), o_cause(D), o_mean(D), o_quant(D), o_cnt(D);
    std::vector<int64_t> n_counts(D, 0);
    for (size_t d = 0; d < D; ++d) {
        const mcdp_propagator &P = *b->props[d];
        if (n[d] < 1) return fail(MCDP_E_INVALID, "statistics need at least one sample per DAG");
        if (!seeds[d] || !mean[d] || (n_q > 0 && !quantiles[d])) return fail(MCDP_E_INVALID, "null argument");
        const bool counts = cause_counts && cause_counts[d];
        n_counts[d] = counts ? P.pred_off.back() + P.E : 0;
        const size_t E = (size_t)P.E, m = (size_t)n[d];
        size_t off = 0;
        auto take = [&](size_t nbytes) {
            const size_t o = off;
            off += align_to(std::max<size_t>(nbytes, 1), a);
            return o;
        };
        o_seed[d] = take(4 * m);
        o_real[d] = take(8 * E * m);
        o_cause[d] = take(counts ? cb * E * m : 0);
        o_mean[d] = take(8 * E);
        o_quant[d] = take(8 * E * (size_t)n_q);
        o_cnt[d] = take(8 * (size_t)n_counts[d]);
        bytes[d] = off;
        if (bytes[d] > (size_t)memory_budget)
            return fail(MCDP_E_NOMEM, "one DAG's realized and cause rows exceed memory_budget");
    }
    // Groups in batch order, each as large as the budget allows; the statistics'
    // per-DAG descriptors follow the largest group in the workspace.
    std::vector<size_t> group_start{0};
    size_t max_group = 0;
    for (size_t d = 0, used = 0; d < D; ++d) {
        if (used + bytes[d] > (size_t)memory_budget) {
            group_start.push_back(d);
            used = 0;
        }
        used += bytes[d];
        max_group = std::max(max_group, used);
    }
    group_start.push_back(D);
    const size_t o_sets = align_to(max_group, a);
    const size_t need = o_sets + align_to(std::max<size_t>(D, 1) * sizeof(RowSet), a);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    if (!b->stats_work || b->stats_bytes < need || b->stats_device != device) {
        if (b->stats_work) {
            (void)hipDeviceSynchronize();
            (void)hipFree(b->stats_work);
            b->stats_work = nullptr;
            b->stats_bytes = 0;
        }
        hipError_t e = hipMalloc(&b->stats_work, need);
        if (e != hipSuccess) {
            b->stats_work = nullptr;
            return fail(MCDP_E_NOMEM, std::string("hipMalloc (batch statistics): ") + hipGetErrorString(e));
        }
        b->stats_bytes = need;
        b->stats_device = device;
    }
    // the walk and then the group's statistics, every DAG's rows in one launch
    // per statistics kernel (small DAGs alone would leave most CUs idle)
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
    auto *base = static_cast<unsigned char *>(b->stats_work);
    auto *sets_dev = reinterpret_cast<RowSet *>(base + o_sets);
    std::vector<RowSet> sets;
    int rc = MCDP_OK;
    std::vector<const int32_t *> sd(D, nullptr);
    std::vector<int64_t> gn(D, 0), gld(D, 1);
    std::vector<unsigned char *> part(D, nullptr);  // a DAG's part of the workspace in its group
    std::vector<double *> real(D, nullptr);
    std::vector<int32_t *> cause(D, nullptr);
    std::vector<uint8_t *> cpos(D, nullptr);
    for (size_t g = 0; rc == MCDP_OK && g + 1 < group_start.size(); ++g) {
        std::fill(gn.begin(), gn.end(), 0);
        std::fill(gld.begin(), gld.end(), 1);
        size_t at = 0;
        for (size_t d = group_start[g]; d < group_start[g + 1]; ++d) {
            unsigned char *p = part[d] = base + at;
            sd[d] = reinterpret_cast<const int32_t *>(p + o_seed[d]);
            real[d] = reinterpret_cast<double *>(p + o_real[d]);
            cause[d] = n_counts[d] && !pos ? reinterpret_cast<int32_t *>(p + o_cause[d]) : nullptr;
            cpos[d] = n_counts[d] && pos ? reinterpret_cast<uint8_t *>(p + o_cause[d]) : nullptr;
            gn[d] = n[d];
            gld[d] = n[d];
            e = hipMemcpyAsync(p + o_seed[d], seeds[d], 4 * (size_t)n[d], hipMemcpyHostToDevice, s);
            if (e != hipSuccess) {
                rc = hip_fail(e, "copy seeds");
                break;
            }
            at += bytes[d];
        }
        if (rc) break;
        // the walk, then (same stream) the statistics
        rc = prepare(b, sd.data(), gn.data(), gld.data(), real.data(), nullptr, pos ? nullptr : cause.data(),
                     pos ? cpos.data() : nullptr);
        if (rc == MCDP_OK) rc = mcdp_batch_launch(b, s);
        sets.clear();
        int64_t rows = 0;
        bool any_counts = false;
        for (size_t d = group_start[g]; rc == MCDP_OK && d < group_start[g + 1]; ++d) {
            unsigned char *p = part[d];
            auto *dm = reinterpret_cast<double *>(p + o_mean[d]);
            auto *dq = n_q > 0 ? reinterpret_cast<double *>(p + o_quant[d]) : nullptr;
            auto *dc = n_counts[d] ? reinterpret_cast<int64_t *>(p + o_cnt[d]) : nullptr;
            RowSet rs;
            rc = row_set(*b->props[d], real[d], cause[d], cpos[d], n[d], n_q, q, dm, dq, dc, rows, rs);
            sets.push_back(rs);
            rows += b->props[d]->E;
            any_counts = any_counts || dc != nullptr;
        }
        if (rc == MCDP_OK &&
            (e = hipMemcpyAsync(sets_dev, sets.data(), sets.size() * sizeof(RowSet), hipMemcpyHostToDevice, s)) !=
                hipSuccess)
            rc = hip_fail(e, "copy statistics descriptors");
        if (rc == MCDP_OK) rc = set_stats(sets_dev, (int32_t)sets.size(), rows, n_q > 0, any_counts, s);
        for (size_t d = group_start[g]; rc == MCDP_OK && d < group_start[g + 1]; ++d) {
            unsigned char *p = part[d];
            const size_t E = (size_t)b->props[d]->E;
            if ((e = hipMemcpyAsync(mean[d], p + o_mean[d], 8 * E, hipMemcpyDeviceToHost, s)) != hipSuccess)
                rc = hip_fail(e, "copy means");
            if (rc == MCDP_OK && n_q > 0 &&
                (e = hipMemcpyAsync(quantiles[d], p + o_quant[d], 8 * E * (size_t)n_q, hipMemcpyDeviceToHost, s)) !=
                    hipSuccess)
                rc = hip_fail(e, "copy quantiles");
            if (rc == MCDP_OK && n_counts[d] &&
                (e = hipMemcpyAsync(cause_counts[d], p + o_cnt[d], 8 * (size_t)n_counts[d], hipMemcpyDeviceToHost,
                                    s)) != hipSuccess)
                rc = hip_fail(e, "copy cause counts");
        }
        // the next group reuses the workspace (and the descriptors' host copy goes away): all of this one is done
        if ((e = hipStreamSynchronize(s)) != hipSuccess && rc == MCDP_OK) rc = hip_fail(e, "hipStreamSynchronize");
        for (size_t d = group_start[g]; d < group_start[g + 1]; ++d) {
            sd[d] = nullptr;
            real[d] = nullptr;
            cause[d] = nullptr;
            cpos[d] = nullptr;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

}  // extern "C"
