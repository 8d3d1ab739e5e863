// Pathwise gradients of the expected weighted lateness (C ABI in
// include/mcdp.h): the reverse plan, its upload, and the entry points.  The
// reference has no derivative of its propagation loop (_core.cpp:332-350);
// this is the loop's adjoint along each event's winning entry, computed over
// the outputs of run_many.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"

using namespace mcdp;

namespace {

constexpr size_t kRevLdsBudget = 150 * 1024;         // LDS per workgroup for slots
constexpr size_t kSlotBytes = (size_t)kLanes * 12;   // adjoint + winner position per sample of a wave
constexpr size_t kPartialBudget = (size_t)1 << 30;   // bytes of partial rows per launch
constexpr size_t kGradWorkspace = (size_t)32 << 30;  // per-sample buffers of mcdp_gradient_host

size_t align_to(size_t x, size_t a) { return (x + a - 1) / a * a; }

// d duration / d minimal duration for link l (see mcdp.h).
uint32_t mul_class(const mcdp_propagator &P, int32_t l, const Record &rec) {
    if (l < 0 || l >= P.L) return kMulNone;
    switch (rec.info & kInfoKindMask) {
        case kNone:
        case kEmpiricalAbs1:
        case kEmpiricalAbs:
            return kMulOne;
        default:
            return P.dmin[(size_t)l] != 0.0 ? kMulRead : kMulDraw;
    }
}

// Waves per workgroup: enough to hold the widest level in one step.
int32_t reverse_waves(const mcdp_propagator &P) {
    size_t width = 0;
    for (const auto &v : P.levels) width = std::max(width, v.size());
    int32_t W = 1;
    while (W < kMaxWaves && (size_t)W * kRevEventsPerWave < width) W *= 2;
    return W;
}

int build_reverse(const mcdp_propagator &P, RevPlan &R) {
    const int32_t E = P.E, L = P.L;
    const int32_t W = reverse_waves(P);
    const int32_t K = kRevEventsPerWave;
    R = RevPlan{};
    R.waves = W;

    // successor lists: (target, position) for every entry that reads its source fresh
    std::vector<std::vector<std::pair<int32_t, int32_t>>> succ((size_t)E);
    for (int32_t e = 0; e < E; ++e) {
        const int64_t a = P.pred_off[(size_t)e], b = P.pred_off[(size_t)e + 1];
        if (b - a > (int64_t)kRevNpredMask) return fail(MCDP_E_INVALID, "too many predecessors for the gradient");
        for (int64_t k = a; k < b; ++k) {
            const PredEntry &pe = P.preds[(size_t)k];
            if (pe.stale) continue;
            succ[(size_t)pe.src].emplace_back(e, (int32_t)(k - a));
        }
    }
    auto cost = [&](int32_t e) {
        return 2 + (int64_t)succ[(size_t)e].size() + 3 * (P.pred_off[(size_t)e + 1] - P.pred_off[(size_t)e]);
    };

    // (a) steps: levels in reverse order, each in as few steps as W x K events allow
    std::vector<int32_t> ev_step((size_t)E, -1);
    R.ev_off.assign(1, 0);
    int32_t step = 0;
    for (auto it = P.levels.rbegin(); it != P.levels.rend(); ++it) {
        const auto &lev = *it;
        if (lev.empty()) continue;
        const size_t cap = (size_t)W * K;
        const size_t nsteps = (lev.size() + cap - 1) / cap;
        for (size_t c = 0; c < nsteps; ++c) {
            const size_t c0 = lev.size() * c / nsteps, c1 = lev.size() * (c + 1) / nsteps;
            std::vector<std::pair<int64_t, int32_t>> items;
            for (size_t i = c0; i < c1; ++i) items.emplace_back(cost(lev[i]), lev[i]);
            std::stable_sort(items.begin(), items.end(), [](auto &x, auto &y) { return x.first > y.first; });
            std::vector<std::vector<int32_t>> per((size_t)W);
            std::vector<int64_t> load((size_t)W, 0);
            for (auto &itm : items) {
                int best = -1;
                for (int w = 0; w < W; ++w)
                    if ((int)per[(size_t)w].size() < K && (best < 0 || load[(size_t)w] < load[(size_t)best])) best = w;
                per[(size_t)best].push_back(itm.second);
                load[(size_t)best] += itm.first;
            }
            for (int w = 0; w < W; ++w) {
                for (int32_t e : per[(size_t)w]) {
                    R.ev_list.push_back(e);
                    ev_step[(size_t)e] = step;
                }
                R.ev_off.push_back((int32_t)R.ev_list.size());
            }
            ++step;
        }
    }
    R.n_steps = step;

    // (b) slots: an event whose adjoint some predecessor gathers holds a slot
    // from its step to the last step of those predecessors (later in reverse).
    std::vector<int32_t> last_use((size_t)E, -1);
    for (int32_t e = 0; e < E; ++e)
        for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k) {
            const PredEntry &pe = P.preds[(size_t)k];
            if (!pe.stale) last_use[(size_t)e] = std::max(last_use[(size_t)e], ev_step[(size_t)pe.src]);
        }
    std::vector<int32_t> slot_of((size_t)E, -1);
    {
        std::vector<std::vector<int32_t>> release((size_t)std::max(R.n_steps, 1));
        std::vector<int32_t> free_slots;
        int32_t used = 0;
        for (int32_t t = 0; t < R.n_steps; ++t) {
            for (int32_t e : release[(size_t)t]) free_slots.push_back(slot_of[(size_t)e]);
            // lowest free slots first, so the busiest slots stay in LDS
            std::sort(free_slots.begin(), free_slots.end(), std::greater<int32_t>());
            for (int32_t i = R.ev_off[(size_t)t * W]; i < R.ev_off[(size_t)(t + 1) * W]; ++i) {
                const int32_t e = R.ev_list[(size_t)i];
                if (last_use[(size_t)e] < 0) continue;
                int32_t s;
                if (!free_slots.empty()) {
                    s = free_slots.back();
                    free_slots.pop_back();
                } else {
                    s = used++;
                }
                slot_of[(size_t)e] = s;
                release[(size_t)last_use[(size_t)e]].push_back(e);
            }
        }
        R.slots = used;
    }
    // two samples per lane when the window of slots fits in LDS that way (and
    // winner positions fit the 16 bits a sample gets until the commit)
    int64_t max_preds = 0;
    for (int32_t e = 0; e < E; ++e) max_preds = std::max<int64_t>(max_preds, P.pred_off[(size_t)e + 1] - P.pred_off[(size_t)e]);
    R.spl = ((size_t)R.slots * kSlotBytes * 2 <= kRevLdsBudget && max_preds <= 32767) ? 2 : 1;
    for (int32_t e : R.ev_list) R.ev_slot.push_back(slot_of[(size_t)e]);

    // (c) records per (step, wave): header, then predecessors; successor pairs
    // apart.  Partial-row columns are numbered in the order a wave writes them
    // (per event: each predecessor entry with a minimal-duration column, plus
    // the source's earliest-time column after a stale one, then the event's
    // earliest-time column), so each wave's writes of a step are one run.
    std::vector<int32_t> unit_row((size_t)L, -1);
    std::vector<std::vector<int32_t>> link_cols((size_t)L), event_cols((size_t)E);
    int32_t next = 0;
    R.step_off.assign(1, 0);
    R.col_off.clear();
    for (size_t g = 0; g + 1 < R.ev_off.size(); ++g) {
        R.col_off.push_back(next);
        for (int32_t i = R.ev_off[g]; i < R.ev_off[g + 1]; ++i) {
            const int32_t e = R.ev_list[(size_t)i];
            const int64_t a = P.pred_off[(size_t)e], b = P.pred_off[(size_t)e + 1];
            Record h{};
            h.a = P.earliest[(size_t)e];
            h.b = P.earliest[(size_t)e] + P.max_delay;
            h.i0 = e;
            h.i1 = slot_of[(size_t)e];
            h.i2 = (int32_t)succ[(size_t)e].size();
            h.info = kInfoHeader | (uint32_t)(b - a);
            R.ev_rec.push_back((int32_t)R.recs.size());
            R.recs.push_back(h);
            std::vector<int32_t> srcs;
            for (int64_t k = a; k < b; ++k) srcs.push_back(P.preds[(size_t)k].src);
            std::sort(srcs.begin(), srcs.end());
            for (int64_t k = a; k < b; ++k) {
                const PredEntry &pe = P.preds[(size_t)k];
                const auto range = std::equal_range(srcs.begin(), srcs.end(), pe.src);
                const int64_t dups = range.second - range.first;
                const Record fwd = link_record(P, pe.edge);
                const uint32_t cls = mul_class(P, pe.edge, fwd);
                Record r = fwd;
                r.a = cls == kMulRead ? 1.0 / P.dmin[(size_t)pe.edge] : 1.0;
                r.i0 = pe.src;
                r.i1 = -1;
                if (cls != kMulNone || pe.stale) {
                    r.i1 = next++;
                    if (cls != kMulNone) link_cols[(size_t)pe.edge].push_back(r.i1);
                    if (pe.stale) event_cols[(size_t)pe.src].push_back(next++);
                }
                r.i2 = pe.edge;
                r.info = (r.info & ~(uint32_t)(kRevClassMask | kRevDup | kRevStale | kInfoOwner | kInfoLast)) |
                         (cls << kRevClassShift) | (dups > 1 ? kRevDup : 0u) | (pe.stale ? kRevStale : 0u);
                if (cls == kMulDraw) {  // its multiplier comes from a row of unit durations
                    int32_t &u = unit_row[(size_t)pe.edge];
                    if (u < 0) {
                        if (R.unit_links.size() > kInfoDistMask) return fail(MCDP_E_INVALID, "too many links with minimal duration 0");
                        u = (int32_t)R.unit_links.size();
                        R.unit_links.push_back(fwd);
                    }
                    r.info = (r.info & ~(kInfoDistMask << kInfoDistShift)) | ((uint32_t)u << kInfoDistShift);
                }
                R.recs.push_back(r);
            }
            event_cols[(size_t)e].insert(event_cols[(size_t)e].begin(), next++);
            if (next > std::numeric_limits<int32_t>::max() - kMaxWaves)
                return fail(MCDP_E_INVALID, "too many gradient columns");
            R.ev_succ.push_back((int32_t)(R.succ.size() / 2));
            for (const auto &sp : succ[(size_t)e]) {
                R.succ.push_back(slot_of[(size_t)sp.first]);
                R.succ.push_back(sp.second);
            }
        }
        R.step_off.push_back((int32_t)R.recs.size());
    }
    if (R.recs.size() > (size_t)std::numeric_limits<int32_t>::max()) return fail(MCDP_E_INVALID, "plan too large");
    // (d) the reduction's map: each output's columns, in writing order
    R.obj_col = next;
    R.n_cols = next + W;
    R.map_off.assign(1, 0);
    for (int32_t l = 0; l < L; ++l) {
        for (int32_t c : link_cols[(size_t)l]) R.map_col.push_back(c);
        R.map_off.push_back((int32_t)R.map_col.size());
    }
    for (int32_t e = 0; e < E; ++e) {
        for (int32_t c : event_cols[(size_t)e]) R.map_col.push_back(c);
        R.map_off.push_back((int32_t)R.map_col.size());
    }
    for (int32_t w = 0; w < W; ++w) R.map_col.push_back(R.obj_col + w);
    R.map_off.push_back((int32_t)R.map_col.size());
    return MCDP_OK;
}

// The reverse plan on the device of `d` (built and uploaded on first use).
int ensure_reverse(mcdp_propagator &P, DeviceCopy &d) {
    if (!P.rev) {
        auto R = std::make_shared<RevPlan>();
        int rc = build_reverse(P, *R);
        if (rc) return rc;
        P.rev = R;
    }
    if (d.rev) return MCDP_OK;
    const RevPlan &R = *P.rev;
    size_t off = 0;
    auto reserve = [&](size_t bytes) {
        const size_t o = off;
        off = align_to(off + bytes, 256);
        return o;
    };
    // per-event blocks, one stream per (step, wave) (mcdp_grad.hpp)
    std::vector<uint32_t> blk;
    std::vector<int32_t> blk_off;
    {
        auto put_f64 = [&](double x) {
            uint64_t u;
            std::memcpy(&u, &x, 8);
            blk.push_back((uint32_t)u);
            blk.push_back((uint32_t)(u >> 32));
        };
        for (size_t g = 0; g + 1 < R.ev_off.size(); ++g) {
            blk_off.push_back((int32_t)blk.size());
            for (int32_t i = R.ev_off[g]; i < R.ev_off[g + 1]; ++i) {
                const Record &h = R.recs[(size_t)R.ev_rec[(size_t)i]];
                const uint32_t n_pred = h.info & kRevNpredMask, n_succ = (uint32_t)h.i2;
                if (n_pred > 0xFFFFu || n_succ > 0xFFFFu)
                    return fail(MCDP_E_INVALID, "too many predecessors or successors of one event for the gradient");
                const size_t start = blk.size();
                blk.push_back((uint32_t)h.i0);
                blk.push_back((uint32_t)h.i1);
                blk.push_back(n_pred | (n_succ << 16));
                blk.push_back(kBlkHeader + kBlkPred * n_pred + kBlkSucc * n_succ);
                put_f64(h.a);
                put_f64(h.b);
                for (uint32_t k = 0; k < n_pred; ++k) {
                    const Record &q = R.recs[(size_t)R.ev_rec[(size_t)i] + 1 + k];
                    put_f64(q.a);
                    blk.push_back((uint32_t)q.i0);
                    blk.push_back((uint32_t)q.i2);
                    blk.push_back(q.info);
                }
                const size_t q0 = (size_t)R.ev_succ[(size_t)i];
                for (uint32_t k = 0; k < n_succ; ++k) {
                    blk.push_back((uint32_t)R.succ[2 * (q0 + k)]);
                    blk.push_back((uint32_t)R.succ[2 * (q0 + k) + 1]);
                }
                if (blk.size() - start != kBlkHeader + kBlkPred * n_pred + kBlkSucc * n_succ)
                    return fail(MCDP_E_INVALID, "internal error: gradient block size");
            }
        }
        blk_off.push_back((int32_t)blk.size());
        for (int k = 0; k < kBlkWindow; ++k) blk.push_back(0u);
        if (blk.size() > (size_t)std::numeric_limits<int32_t>::max()) return fail(MCDP_E_INVALID, "plan too large");
    }
    const size_t o_blk = reserve(blk.size() * 4), o_boff = reserve(blk_off.size() * 4);
    const size_t o_evoff = reserve(R.ev_off.size() * 4);
    // the kernel reads a (step, wave)'s slots as one row of kRevEventsPerWave: padding after the last
    const size_t o_evs = reserve((R.ev_slot.size() + kRevEventsPerWave) * 4), o_moff = reserve(R.map_off.size() * 4);
    const size_t o_unit = reserve(R.unit_links.size() * sizeof(Record));
    const size_t o_coloff = reserve(R.col_off.size() * 4);
    const size_t o_mcol = reserve(R.map_col.size() * 4);
    const size_t total = std::max<size_t>(off, 256);
    std::vector<unsigned char> host(total, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) std::memcpy(host.data() + o, src, bytes);
    };
    put(o_blk, blk.data(), blk.size() * 4);
    put(o_boff, blk_off.data(), blk_off.size() * 4);
    put(o_evoff, R.ev_off.data(), R.ev_off.size() * 4);
    put(o_evs, R.ev_slot.data(), R.ev_slot.size() * 4);
    put(o_unit, R.unit_links.data(), R.unit_links.size() * sizeof(Record));
    put(o_coloff, R.col_off.data(), R.col_off.size() * 4);
    put(o_moff, R.map_off.data(), R.map_off.size() * 4);
    put(o_mcol, R.map_col.data(), R.map_col.size() * 4);
    void *blob = nullptr;
    hipError_t e = hipMalloc(&blob, total);
    if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (reverse plan): ") + hipGetErrorString(e));
    e = hipMemcpy(blob, host.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(blob);
        return hip_fail(e, "hipMemcpy (reverse plan)");
    }
    auto *b = static_cast<unsigned char *>(blob);
    d.rev = blob;
    RevParams &p = d.rev_base;
    p = RevParams{};
    p.blocks = reinterpret_cast<const uint32_t *>(b + o_blk);
    p.blk_off = reinterpret_cast<const int32_t *>(b + o_boff);
    p.ev_off = reinterpret_cast<const int32_t *>(b + o_evoff);
    p.ev_slot = reinterpret_cast<const int32_t *>(b + o_evs);
    p.col_off = reinterpret_cast<const int32_t *>(b + o_coloff);
    p.dists = d.base.dists;
    p.alias = d.base.alias;
    p.earliest = d.base.earliest;
    p.n_steps = R.n_steps;
    p.waves = R.waves;
    p.lds_slots = std::min<int32_t>(R.slots, (int32_t)(kRevLdsBudget / (kSlotBytes * R.spl)));
    p.global_slots = R.slots - p.lds_slots;
    p.L = P.L;
    p.n_cols = R.n_cols;
    p.obj_col = R.obj_col;
    d.rev_map_off = reinterpret_cast<const int32_t *>(b + o_moff);
    d.rev_map_col = reinterpret_cast<const int32_t *>(b + o_mcol);
    d.rev_unit_links = reinterpret_cast<const Record *>(b + o_unit);
    return MCDP_OK;
}

// Grows the propagator's gradient workspace on `d` to `bytes`.
int ensure_workspace(DeviceCopy &d, size_t bytes) {
    if (d.rev_work && d.rev_work_bytes >= bytes) return MCDP_OK;
    if (d.rev_work) {
        (void)hipDeviceSynchronize();  // earlier launches may still use it
        (void)hipFree(d.rev_work);
        d.rev_work = nullptr;
        d.rev_work_bytes = 0;
    }
    hipError_t e = hipMalloc(&d.rev_work, bytes);
    if (e != hipSuccess) {
        d.rev_work = nullptr;
        return fail(MCDP_E_NOMEM, std::string("hipMalloc (gradient workspace): ") + hipGetErrorString(e));
    }
    d.rev_work_bytes = bytes;
    return MCDP_OK;
}

}  // namespace

int mcdp::reverse_run(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev,
                      const double *realized_dev, const double *durations_dev, const int32_t *cause_dev, int64_t n,
                      int64_t ld, int accumulate, double *objective_dev, double *grad_dmin_dev,
                      double *grad_earliest_dev, void *given, size_t given_bytes, hipStream_t s) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n < 0 || (n > 0 && ld < n)) return fail(MCDP_E_INVALID, "need 0 <= n <= ld");
    if (n > std::numeric_limits<int32_t>::max() - 256 || ld > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "n and ld must be below 2^31 - 256 per call");
    if (!objective_dev || (!grad_dmin_dev && prop->L > 0) || (!grad_earliest_dev && prop->E > 0))
        return fail(MCDP_E_INVALID, "gradient outputs are null");
    if (n > 0 && prop->E > 0 && (!weights_dev || !realized_dev || !cause_dev || (!durations_dev && prop->L > 0)))
        return fail(MCDP_E_INVALID, "weights, realized, durations and cause are required");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: gradients run on the GPU only");
    std::lock_guard<std::mutex> lock(prop->mu);
    mcdp_propagator &P = *prop;
    DeviceCopy *d = nullptr;
    int rc = MCDP_OK;
    if (given) {
        // nothing may be allocated: both plans must be on this device already
        int device = 0;
        HIP_TRY(hipGetDevice(&device), "hipGetDevice");
        for (auto &c : P.dev)
            if (c.device == device) d = &c;
        if (!d || !d->rev || !P.rev)
            return fail(MCDP_E_INVALID, "the propagator's plans are not resident on this device: call "
                                        "mcdp_reverse_device on it once first (n = 0 is enough)");
    } else {
        rc = ensure_device(P, d);
        if (rc) return rc;
        rc = ensure_reverse(P, *d);
        if (rc) return rc;
    }
    const int64_t n_unit = (int64_t)P.rev->unit_links.size();
    if (n > 0 && n_unit > 0 && !seeds_dev)
        return fail(MCDP_E_INVALID, "seeds are required: a relative link has minimal duration 0");
    const RevParams &base = d->rev_base;
    const int spl = P.rev->spl;
    const int64_t tile = (int64_t)kLanes * spl;
    const int64_t tiles = (n + tile - 1) / tile;
    const size_t row_bytes = (size_t)base.n_cols * sizeof(double);
    // launches cover whole groups of the column reduction, so sums do not depend on the split
    int64_t max_tiles = std::max<int64_t>(kReduceGroup, (int64_t)(kPartialBudget / row_bytes));
    if (P.rev_max_tiles > 0) max_tiles = std::min(max_tiles, P.rev_max_tiles);
    max_tiles = std::min<int64_t>(max_tiles, (int64_t)kReduceGroup * 65535);
    max_tiles = std::max<int64_t>(kReduceGroup, max_tiles / kReduceGroup * kReduceGroup);
    const size_t lds = (size_t)base.lds_slots * kSlotBytes * spl;
    int cus = 0, bpc = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device), "hipDeviceGetAttribute");
    HIP_TRY(reverse_occupancy(base.waves, lds, base.global_slots > 0, spl, &bpc),
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (bpc < 1) return fail(MCDP_E_DEVICE, "reverse kernel cannot be resident with " + std::to_string(lds) + " bytes of LDS");
    const int64_t resident = (int64_t)cus * bpc;
    int64_t chunk_tiles = std::min<int64_t>(std::max<int64_t>(tiles, 1), max_tiles);
    if (n_unit > 0) {  // unit-duration rows for the launch's samples stay within the same budget
        const int64_t cap = (int64_t)(kPartialBudget / ((size_t)n_unit * tile * 8));
        if (chunk_tiles > cap)
            chunk_tiles = std::max<int64_t>(kReduceGroup, cap / kReduceGroup * kReduceGroup);
    }
    // Workspace of a launch of ct tiles: partial rows, global slots, unit
    // rows, the reduction's group sums.
    struct Layout {
        size_t o_gslots, o_unit, o_groups, total;
        int32_t ld_unit;
    };
    auto layout = [&](int64_t ct) {
        Layout w;
        const int64_t grid_max = std::min<int64_t>(ct, resident);
        w.o_gslots = align_to(row_bytes * (size_t)ct, 256);
        const size_t gslot_bytes = (size_t)grid_max * base.global_slots * tile * 2 * sizeof(double);
        w.o_unit = w.o_gslots + align_to(std::max<size_t>(gslot_bytes, 256), 256);
        w.ld_unit = (int32_t)(ct * tile);
        w.o_groups = w.o_unit + align_to(std::max<size_t>((size_t)n_unit * (size_t)w.ld_unit * 8, 256), 256);
        const size_t group_bytes = row_bytes * (size_t)((ct + kReduceGroup - 1) / kReduceGroup);
        w.total = w.o_groups + std::max<size_t>(group_bytes, 256);
        return w;
    };
    Layout w = layout(chunk_tiles);
    void *work_base = nullptr;
    if (given) {
        // the largest launch (whole reduction groups) that fits the given workspace
        while (w.total > given_bytes && chunk_tiles > kReduceGroup) {
            chunk_tiles = std::max<int64_t>(kReduceGroup, (chunk_tiles - 1) / kReduceGroup * kReduceGroup);
            w = layout(chunk_tiles);
        }
        if (w.total > given_bytes)
            return fail(MCDP_E_NOMEM, "the reverse walk of this DAG needs " + std::to_string(w.total) +
                                          " bytes of workspace per launch; the fixed scratch holds " +
                                          std::to_string(given_bytes) + " bytes");
        work_base = given;
    } else {
        rc = ensure_workspace(*d, w.total);
        if (rc) return rc;
        work_base = d->rev_work;
    }
    auto *work = static_cast<unsigned char *>(work_base);
    RevParams p = base;
    p.weights = weights_dev;
    p.realized = realized_dev;
    p.durations = durations_dev;
    p.cause = cause_dev;
    p.n = (int32_t)n;
    p.ld = (int32_t)std::max<int64_t>(ld, 1);
    p.partial = reinterpret_cast<double *>(work);
    p.gslots = reinterpret_cast<double *>(work + w.o_gslots);
    double *unit = reinterpret_cast<double *>(work + w.o_unit);
    double *groups = reinterpret_cast<double *>(work + w.o_groups);
    p.unit = unit;
    p.ld_unit = w.ld_unit;
    if (tiles == 0 || P.E == 0) {  // no samples: sums are 0 (or unchanged)
        HIP_TRY(launch_reduce_columns(p.partial, 0, p.n_cols, groups, d->rev_map_off, d->rev_map_col, P.L, P.E,
                                      accumulate, grad_dmin_dev, grad_earliest_dev, objective_dev, s),
                "reduce kernel launch");
        return MCDP_OK;
    }
    for (int64_t t0 = 0; t0 < tiles; t0 += chunk_tiles) {
        p.tile0 = (int32_t)t0;
        p.n_tiles = (int32_t)std::min<int64_t>(chunk_tiles, tiles - t0);
        const int grid = (int)std::min<int64_t>(p.n_tiles, resident);
        HIP_TRY(launch_unit_draws(d->rev_unit_links, (int32_t)n_unit, seeds_dev, (int32_t)(t0 * tile),
                                  (int32_t)(p.n_tiles * tile), (int32_t)n, base.dists, base.alias, unit, w.ld_unit, s),
                "unit draw kernel launch");
        HIP_TRY(launch_reverse(p, spl, grid, lds, s), "reverse kernel launch");
        HIP_TRY(launch_reduce_columns(p.partial, p.n_tiles, p.n_cols, groups, d->rev_map_off, d->rev_map_col, P.L, P.E,
                                      accumulate || t0 > 0, grad_dmin_dev, grad_earliest_dev, objective_dev, s),
                "reduce kernel launch");
    }
    return MCDP_OK;
}

extern "C" {

int mcdp_propagator_set_times(mcdp_propagator *prop, const double *dmin, const double *earliest) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    std::lock_guard<std::mutex> lock(prop->mu);
    mcdp_propagator &P = *prop;
    if (dmin) std::copy(dmin, dmin + P.L, P.dmin.begin());
    if (earliest) std::copy(earliest, earliest + P.E, P.earliest.begin());
    for (Record &r : P.recs) {
        if (r.info & kInfoHeader) {
            r.a = P.earliest[(size_t)r.i0];
            r.b = P.earliest[(size_t)r.i0] + P.max_delay;
        } else if (r.info & kInfoEdgeValid) {
            r.a = P.dmin[(size_t)r.i2];
        }
    }
    for (Record &r : P.orph) r.a = P.dmin[(size_t)r.i2];
    P.rev.reset();
    P.drop_device_copies();
    return MCDP_OK;
}

int mcdp_propagator_get_times(const mcdp_propagator *prop, double *dmin, double *earliest) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (dmin) std::copy(prop->dmin.begin(), prop->dmin.end(), dmin);
    if (earliest) std::copy(prop->earliest.begin(), prop->earliest.end(), earliest);
    return MCDP_OK;
}

int mcdp_reverse_device(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev,
                        const double *realized_dev, const double *durations_dev, const int32_t *cause_dev, int64_t n,
                        int64_t ld, int accumulate, double *objective_dev, double *grad_dmin_dev,
                        double *grad_earliest_dev, void *stream) {
    return reverse_run(prop, seeds_dev, weights_dev, realized_dev, durations_dev, cause_dev, n, ld, accumulate,
                       objective_dev, grad_dmin_dev, grad_earliest_dev, nullptr, 0, static_cast<hipStream_t>(stream));
}

int mcdp_propagator_gradient_info(mcdp_propagator *prop, int64_t *out, int n_out) {
    if (!prop || !out) return fail(MCDP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(prop->mu);
    if (!prop->rev) {
        auto R = std::make_shared<RevPlan>();
        int rc = build_reverse(*prop, *R);
        if (rc) return rc;
        prop->rev = R;
    }
    const RevPlan &R = *prop->rev;
    const int64_t v[6] = {R.n_steps, R.waves, R.slots,
                          std::min<int64_t>(R.slots, (int64_t)(kRevLdsBudget / (kSlotBytes * R.spl))), R.n_cols, R.spl};
    for (int i = 0; i < n_out && i < 6; ++i) out[i] = v[i];
    return MCDP_OK;
}

int mcdp_debug_reverse_tiles(mcdp_propagator *prop, int64_t max_tiles) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (max_tiles < 0) return fail(MCDP_E_INVALID, "max_tiles must be >= 0");
    std::lock_guard<std::mutex> lock(prop->mu);
    prop->rev_max_tiles = max_tiles;
    return MCDP_OK;
}

int mcdp_gradient_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, const double *weights, int64_t chunk,
                       double *objective, double *grad_dmin, double *grad_earliest) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n < 1 || !seeds) return fail(MCDP_E_INVALID, "a gradient needs at least one seed");
    if (!weights || !objective || (!grad_dmin && prop->L > 0) || (!grad_earliest && prop->E > 0))
        return fail(MCDP_E_INVALID, "null argument");
    if (chunk < 0) return fail(MCDP_E_INVALID, "chunk must be >= 0");
    const int64_t E = prop->E, L = prop->L;
    const int64_t per_sample = 12 * E + 8 * L + 4;
    int64_t tile = kLanes;
    {
        std::lock_guard<std::mutex> lock(prop->mu);
        if (!prop->rev) {
            auto R = std::make_shared<RevPlan>();
            int rc = build_reverse(*prop, *R);
            if (rc) return rc;
            prop->rev = R;
        }
        tile = (int64_t)kLanes * prop->rev->spl;
    }
    if (chunk == 0) chunk = std::max<int64_t>(tile, (int64_t)(kGradWorkspace / (size_t)std::max<int64_t>(per_sample, 1)));
    chunk = std::min<int64_t>(chunk, std::numeric_limits<int32_t>::max() - 256);
    // whole groups of the column reduction: sums do not depend on the chunking
    const int64_t unit = tile * kReduceGroup;
    chunk = std::max<int64_t>(unit, chunk / unit * unit);
    chunk = std::min<int64_t>(chunk, (n + unit - 1) / unit * unit);
    const size_t o_seeds = 0, o_real = align_to((size_t)chunk * 4, 256);
    const size_t o_dur = o_real + align_to((size_t)(E * chunk) * 8, 256);
    const size_t o_cause = o_dur + align_to((size_t)(L * chunk) * 8, 256);
    const size_t o_w = o_cause + align_to((size_t)(E * chunk) * 4, 256);
    const size_t o_out = o_w + align_to((size_t)E * 8, 256);
    const size_t bytes = o_out + (size_t)(1 + L + E) * 8;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: gradients run on the GPU only");
    std::lock_guard<std::mutex> host_lock(prop->host_mu);
    void *buf = nullptr;
    hipStream_t st = nullptr;
    int rc = host_workspace(*prop, bytes, buf, st);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    auto *b = static_cast<unsigned char *>(buf);
    auto *d_seeds = reinterpret_cast<int32_t *>(b + o_seeds);
    auto *d_real = reinterpret_cast<double *>(b + o_real);
    auto *d_dur = reinterpret_cast<double *>(b + o_dur);
    auto *d_cause = reinterpret_cast<int32_t *>(b + o_cause);
    auto *d_w = reinterpret_cast<double *>(b + o_w);
    auto *d_obj = reinterpret_cast<double *>(b + o_out);
    double *d_gd = d_obj + 1, *d_ge = d_obj + 1 + L;
    if (E > 0 && (e = hipMemcpyAsync(d_w, weights, (size_t)E * 8, hipMemcpyHostToDevice, st)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy");
    for (int64_t c0 = 0; c0 < n && !rc; c0 += chunk) {
        const int64_t nc = std::min<int64_t>(chunk, n - c0);
        if ((e = hipMemcpyAsync(d_seeds, seeds + c0, (size_t)nc * 4, hipMemcpyHostToDevice, st)) != hipSuccess) {
            rc = hip_fail(e, "hipMemcpy");
            break;
        }
        rc = mcdp_run_many_device(prop, d_seeds, nc, nc, d_real, d_dur, d_cause, st);
        if (!rc)
            rc = mcdp_reverse_device(prop, d_seeds, d_w, d_real, d_dur, d_cause, nc, nc, c0 > 0, d_obj, d_gd, d_ge,
                                     st);
        // the next chunk's seeds overwrite this chunk's (the copy is asynchronous for pageable memory)
        if (!rc && (e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_fail(e, "gradient");
    }
    std::vector<double> out((size_t)(1 + L + E));
    if (!rc && (e = hipMemcpyAsync(out.data(), d_obj, out.size() * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess && !rc) rc = hip_fail(e, "gradient");
    host_workspace_trim(*prop);
    if (rc) return rc;
    const double dn = (double)n;
    objective[0] = out[0] / dn;
    for (int64_t l = 0; l < L; ++l) grad_dmin[l] = out[(size_t)(1 + l)] / dn;
    for (int64_t i = 0; i < E; ++i) grad_earliest[i] = out[(size_t)(1 + L + i)] / dn;
    return MCDP_OK;
}

}  // extern "C"
