// gfx950 kernels of the Monte Carlo DAG propagation path.
//
// walk_kernel fuses the two hot loops of Simulator::run (_core.cpp:323-329
// delay sampling, :332-350 max-plus propagation) for a tile of 64 x kSpl
// samples per workgroup: lane l of every wave carries samples
// tile * 64 kSpl + i * 64 + l, i < kSpl.
//   * plan records (event headers; predecessors with their distribution
//     inline) are wave-uniform and arrive through scalar loads, the next one
//     prefetched while the current one is processed; the distribution switch
//     is a scalar branch taken once for the lane's kSpl samples, whose draws
//     then run as independent instruction chains;
//   * the W waves of a workgroup split each step's events (a step is part of
//     one topological level); results stay in registers until a barrier, then
//     go to LDS slots that later steps read;
//   * LDS is slot-major, a slot holding 64 x kSpl doubles with a lane's
//     samples adjacent, so a lane reads or writes a slot with one ds_read_b64
//     or ds_read_b128;
//   * outputs are sample-contiguous rows, so each store instruction of a wave
//     writes a 512-byte run (nontemporal: nothing reads them back).
// The arithmetic of propagation is FP64 add and compare-select only, in the
// reference's order: t = realized[src] + dur; t = (upper < t) ? upper : t;
// if (t >= best) {best = t; cause = src}; realized = (upper < best) ? upper : best.
// std::min is (b < a) ? b : a, so a NaN candidate is never taken; no v_min_f64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mcdp_kernels.hpp"
#include "mcdp_sampler.hpp"

namespace mcdp {

// A lane's kSpl samples of one LDS slot.
template <int kSpl>
__device__ __forceinline__ void lds_load(const double *lds, int slot, int lane, double (&v)[kSpl]) {
    if constexpr (kSpl == 1) {
        v[0] = lds[slot * kLanes + lane];
    } else {
        static_assert(kSpl == 2, "one or two samples per lane");
        const double2 t = reinterpret_cast<const double2 *>(lds)[slot * kLanes + lane];
        v[0] = t.x;
        v[1] = t.y;
    }
}
template <int kSpl>
__device__ __forceinline__ void lds_store(double *lds, int slot, int lane, const double (&v)[kSpl]) {
    if constexpr (kSpl == 1) {
        lds[slot * kLanes + lane] = v[0];
    } else {
        reinterpret_cast<double2 *>(lds)[slot * kLanes + lane] = double2{v[0], v[1]};
    }
}

// Element offset of (row, column 0) in an array with leading dimension ld:
// a 32 x 32 -> 64-bit product (rows and ld are non-negative and < 2^31).
__device__ __forceinline__ uint64_t row_offset(int32_t row, uint32_t ld) { return (uint64_t)(uint32_t)row * ld; }

// Issue priority from the share of its step's records a wave has left: 3 in
// the first quarter down to 0 in the last.  A SIMD issues its oldest waves
// first at equal priority, so with equal work its youngest wave would finish
// every step last while the others wait at the step's barrier.  A wave later
// on its SIMD (waves go to the four SIMDs in turn: rank wave / 4) counts
// itself a further 1/16 of the step behind per rank.
__device__ __forceinline__ void step_priority(int left, int total, int wave) {
    const int l = left + (((wave >> 2) * total) >> 4);
    if (4 * l > 3 * total)
        __builtin_amdgcn_s_setprio(3);
    else if (2 * l > total)
        __builtin_amdgcn_s_setprio(2);
    else if (4 * l > total)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

// Alias tables staged in LDS after the slots (fast sampled variant), so an
// empirical draw is a ds_read rather than a vector memory load.
template <bool kInjected, int kPaths, int kSpl>
__device__ __forceinline__ AliasRec *lds_alias_of(const WalkParams &p, double *lds) {
    return reinterpret_cast<AliasRec *>(lds + (size_t)p.slots * kLanes * kSpl);
}
template <bool kInjected, int kPaths, int kSpl>
__device__ __forceinline__ void stage_alias(const WalkParams &p, double *lds) {
    if (!kInjected) {
        if (kPaths == kPathsFast) {
            AliasRec *const dst = lds_alias_of<kInjected, kPaths, kSpl>(p, lds);
            const auto src = as_const_space(p.alias);
            for (int i = threadIdx.x; i < p.n_alias; i += blockDim.x) dst[i] = src[i];
        }
        __syncthreads();
    }
}

// One tile of 64 x kSpl samples of one plan: every step, then the orphan
// activities.  Every wave of the workgroup calls it; it ends with all waves
// past the last step's barrier.
// The LDS slots of one (step, wave)'s events.
template <int K>
struct SlotRow {
    int32_t v[K];
};

// kPos: the cause output is cause_pos (the winning entry's position), not cause.
template <bool kInjected, int kPaths, int K, int kSpl, bool kPos = false>
__device__ __forceinline__ void walk_tile(const WalkParams &p, int32_t tile, double *lds, const LogTable &lt) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = p.waves;
    const uint32_t ld = (uint32_t)p.ld;
    // records that write their duration row (none when no durations output)
    const uint32_t owner_bit = p.durations ? (uint32_t)kInfoOwner : 0u;
    // The plan never changes during a launch: reading it through the constant
    // address space lets wave-uniform record loads become scalar loads.
    const auto step_off = as_const_space(p.step_off);
    const auto recs = as_const_space(p.recs);
    const auto step_slots = as_const_space(reinterpret_cast<const SlotRow<K> *>(p.step_slots));
    const auto dists = as_const_space(p.dists);
    const auto earliest = as_const_space(p.earliest);

    AliasRec *const lds_alias = lds_alias_of<kInjected, kPaths, kSpl>(p, lds);
    auto draw = [&](const Record &r, const uint32_t(&seed)[kSpl], double(&dur)[kSpl]) {
        if (kPaths == kPathsFast)
            draw_durations<kSpl>(r.info, r.a, r.b, (uint32_t)r.i2, seed, lds_alias, dists, lt, dur);
        else
            draw_durations<kSpl>(r.info, r.a, r.b, (uint32_t)r.i2, seed, as_const_space(p.alias), dists, lt, dur);
    };

    // Lanes past the last sample (last tile only) take over sample n - 1:
    // same seed or injected row, so they compute its values bit for bit and
    // their stores rewrite its bytes unchanged.  No store needs a lane mask.
    uint32_t s[kSpl];
    uint32_t seed[kSpl];
#pragma unroll
    for (int i = 0; i < kSpl; ++i) {
        const int32_t si = tile * (kLanes * kSpl) + i * kLanes + lane;
        s[i] = si < p.n ? si : p.n - 1;
        seed[i] = kInjected ? 0u : (uint32_t)p.seeds[s[i]];
    }

    for (int t = 0; t < p.n_steps; ++t) {
        const int r0 = step_off[t * W + wave];
        const int r1 = step_off[t * W + wave + 1];
        double res[K][kSpl];
        int j = 0;
        if (r0 < r1) {
            // One stream of records: an event header, then its predecessors
            // in input order.  The stream ends in a padding record, so the
            // one-ahead prefetch never leaves the array.
            Record cur = recs[r0];
            double best[kSpl], upper = 0.0;
            int32_t who[kSpl], event = 0;
            int hdr = r0;  // kPos: the current event's header record (entry k is record hdr + 1 + k)
#pragma unroll
            for (int i = 0; i < kSpl; ++i) {
                best[i] = 0.0;
                who[i] = -1;
            }
            for (int r = r0; r < r1; ++r) {
                const Record nxt = recs[r + 1];
                if (cur.info & kInfoHeader) {
                    step_priority(r1 - r, r1 - r0, wave);
                    upper = cur.b;
                    // one vector copy per event: the caps and the final clamp
                    // select against it without rebuilding it from SGPRs
                    asm volatile("" : "+v"(upper));
                    event = cur.i0;
                    if (kPos) hdr = r;
#pragma unroll
                    for (int i = 0; i < kSpl; ++i) {
                        best[i] = cur.a;
                        who[i] = -1;
                    }
                } else {
                    double dur[kSpl];
                    if (kInjected) {
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) dur[i] = 0.0;
                        if (cur.info & kInfoEdgeValid) {
                            const uint64_t row = row_offset(cur.i2, ld);
#pragma unroll
                            for (int i = 0; i < kSpl; ++i)
                                dur[i] = __builtin_nontemporal_load(p.dur_in + row + s[i]);
                        }
                    } else {
                        // an edge outside the link table is kind kNone with
                        // minimal duration 0.0, and owns no row
                        draw(cur, seed, dur);
                        if (cur.info & owner_bit) {
                            const uint64_t row = row_offset(cur.i2, ld);
#pragma unroll
                            for (int i = 0; i < kSpl; ++i)
                                __builtin_nontemporal_store(dur[i], p.durations + row + s[i]);
                        }
                    }
                    double rs[kSpl];
                    if (cur.i1 >= 0) {
                        lds_load<kSpl>(lds, cur.i1, lane, rs);
                    } else if (kPaths == kPathsGeneral && cur.i1 == kSrcGlobal) {
                        const uint64_t row = row_offset(cur.i0, ld);
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) rs[i] = p.realized[row + s[i]];
                    } else {
                        const double e = earliest[cur.i0];
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) rs[i] = e;
                    }
                    // the cause: the source event, or the entry's position
                    const int32_t tag = kPos ? r - hdr - 1 : cur.i0;
#pragma unroll
                    for (int i = 0; i < kSpl; ++i) {
                        double tt = rs[i] + dur[i];
                        tt = (upper < tt) ? upper : tt;
                        if (tt >= best[i]) {
                            best[i] = tt;
                            who[i] = tag;
                        }
                    }
                }
                if (cur.info & kInfoLast) {  // the event's last predecessor (or a header with none)
                    const uint64_t row = row_offset(event, ld);
#pragma unroll
                    for (int i = 0; i < kSpl; ++i) {
                        const double rr = (upper < best[i]) ? upper : best[i];
                        __builtin_nontemporal_store(rr, p.realized + row + s[i]);
                        if (kPos) {
                            if (p.cause_pos) __builtin_nontemporal_store((uint8_t)who[i], p.cause_pos + row + s[i]);
                        } else if (p.cause) {
                            __builtin_nontemporal_store(who[i], p.cause + row + s[i]);
                        }
                        res[j][i] = rr;
                    }
                    ++j;
                }
                cur = nxt;
            }
        }
        // The step's slots in one scalar load, in flight across the barrier.
        const SlotRow<K> srow = step_slots[t * W + wave];
        // Every read of this step is done before any slot is overwritten,
        // and every commit is visible before the next step reads it.
        __syncthreads();
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int slot = srow.v[i];
            if (i < j && slot >= 0) lds_store<kSpl>(lds, slot, lane, res[i]);
        }
        __syncthreads();
    }

    // Activities no predecessor entry references still get a duration row.
    if (!kInjected && p.durations) {
        const auto orph = as_const_space(p.orph);
        for (int o = wave; o < p.n_orphans; o += W) {
            const Record rec = orph[o];
            double dur[kSpl];
            draw(rec, seed, dur);
            const uint64_t row = row_offset(rec.i2, ld);
#pragma unroll
            for (int i = 0; i < kSpl; ++i)
                __builtin_nontemporal_store(dur[i], p.durations + row + s[i]);
        }
    }
}

template <bool kInjected, int kPaths, int K, int kSpl, bool kPos = false>
__global__ void __launch_bounds__(kMaxWaves * kLanes) walk_kernel(WalkParams p) {
    extern __shared__ double lds[];
    __shared__ LogTableLds log_lds;
    if (!kInjected) stage_log_table(log_lds);
    stage_alias<kInjected, kPaths, kSpl>(p, lds);  // its barrier also publishes the log table
    for (int32_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x)
        walk_tile<kInjected, kPaths, K, kSpl, kPos>(p, tile, lds, log_lds.view());
}

// Several plans in one launch: workgroups take (plan, tile) items from a
// queue (largest plans first, as ordered by the host) until it is empty;
// every workgroup leaves through the same check of the queue head.
template <bool kInjected, int kPaths, int K, int kSpl, bool kPos = false>
__global__ void __launch_bounds__(kMaxWaves * kLanes) walk_batch_kernel(WalkBatchParams b) {
    extern __shared__ double lds[];
    __shared__ int32_t next_item;
    __shared__ LogTableLds log_lds;
    stage_log_table(log_lds);  // published by the loop's first barrier
    int32_t staged = -1;
    const auto item_dag = as_const_space(b.item_dag);
    const auto item_tile = as_const_space(b.item_tile);
    const auto params = as_const_space(b.params);
    for (;;) {
        __syncthreads();  // the previous tile is done with LDS and next_item
        if (threadIdx.x == 0) next_item = atomicAdd(b.counter, 1);
        __syncthreads();
        const int32_t item = __builtin_amdgcn_readfirstlane(next_item);
        if (item >= b.n_items) break;
        const int32_t d = item_dag[item];
        const WalkParams p = params[d];
        if (d != staged) {
            stage_alias<kInjected, kPaths, kSpl>(p, lds);
            staged = d;
        }
        walk_tile<kInjected, kPaths, K, kSpl, kPos>(p, item_tile[item], lds, log_lds.view());
    }
}

// One step of the walk for a few samples, spread over the device instead of
// over one workgroup's waves: workgroup (i, g) is one wave evaluating event
// e0 + i of the schedule for lane group g (samples 64 g + lane).  A
// predecessor's realized time comes from its output row, written by the
// launch of an earlier step (or is its earliest time, when stale), so there
// are no slots and no barriers; the stream orders the steps.  Per sample the
// arithmetic is walk_tile's: the same records, draws and operation order.
template <bool kInjected>
__global__ void __launch_bounds__(kLanes) walk_wide_kernel(WalkParams p, int32_t e0) {
    const int32_t si = (int32_t)blockIdx.y * kLanes + (int32_t)threadIdx.x;
    const int32_t s = si < p.n ? si : p.n - 1;
    const uint32_t ld = (uint32_t)p.ld;
    const auto recs = as_const_space(p.recs);
    const auto earliest = as_const_space(p.earliest);
    const LogTable lt = global_log_table();
    const uint32_t seed[1] = {kInjected ? 0u : (uint32_t)p.seeds[s]};
    const int32_t r0 = as_const_space(p.ev_rec)[e0 + (int32_t)blockIdx.x];
    // The header and the first batch of records load together.
    const Record h = recs[r0];
    Record c[kWideBatch];
#pragma unroll
    for (int u = 0; u < kWideBatch; ++u) c[u] = recs[r0 + 1 + u];
    const double upper = h.b;
    double best = h.a;
    int32_t who = -1;
    // Predecessors in batches: the batch's records (scalar loads), then its
    // predecessor times and injected durations (vector loads), all in flight
    // together; then the draws and the max-plus steps in input order.  The
    // record array ends in kWideBatch padding records, so a batch never
    // reads past it.
    for (int32_t q0 = 0; q0 < h.i2; q0 += kWideBatch) {
        double rs[kWideBatch], dur[kWideBatch];
        if (q0 > 0) {
#pragma unroll
            for (int u = 0; u < kWideBatch; ++u) c[u] = recs[r0 + 1 + q0 + u];
        }
#pragma unroll
        for (int u = 0; u < kWideBatch; ++u) {
            rs[u] = 0.0;
            dur[u] = 0.0;
            if (q0 + u < h.i2) {
                rs[u] = c[u].i1 == kSrcStale ? earliest[c[u].i0] : p.realized[row_offset(c[u].i0, ld) + s];
                if (kInjected && (c[u].info & kInfoEdgeValid)) dur[u] = p.dur_in[row_offset(c[u].i2, ld) + s];
            }
        }
#pragma unroll
        for (int u = 0; u < kWideBatch; ++u) {
            if (q0 + u >= h.i2) break;
            if (!kInjected) {
                double d1[1];
                draw_durations<1>(c[u].info, c[u].a, c[u].b, (uint32_t)c[u].i2, seed, as_const_space(p.alias),
                                  as_const_space(p.dists), lt, d1);
                dur[u] = d1[0];
                if (p.durations && (c[u].info & kInfoOwner))
                    __builtin_nontemporal_store(dur[u], p.durations + row_offset(c[u].i2, ld) + s);
            }
            double tt = rs[u] + dur[u];
            tt = (upper < tt) ? upper : tt;
            if (tt >= best) {
                best = tt;
                who = p.cause_pos ? q0 + u : c[u].i0;  // the entry's position, or the source event
            }
        }
    }
    // realized is read by later steps: an ordinary store; cause is not
    const uint64_t row = row_offset(h.i0, ld);
    p.realized[row + s] = (upper < best) ? upper : best;
    if (p.cause_pos)
        __builtin_nontemporal_store((uint8_t)who, p.cause_pos + row + s);
    else if (p.cause)
        __builtin_nontemporal_store(who, p.cause + row + s);
}

// Durations of the activities no predecessor entry references, for the wide
// steps' samples: workgroup (o, g) draws orphan o for lane group g.
__global__ void __launch_bounds__(kLanes) walk_wide_orphans_kernel(WalkParams p) {
    const int32_t si = (int32_t)blockIdx.y * kLanes + (int32_t)threadIdx.x;
    const int32_t s = si < p.n ? si : p.n - 1;
    const Record rec = as_const_space(p.orph)[blockIdx.x];
    const uint32_t seed[1] = {(uint32_t)p.seeds[s]};
    double dur[1];
    draw_durations<1>(rec.info, rec.a, rec.b, (uint32_t)rec.i2, seed, as_const_space(p.alias), as_const_space(p.dists),
                      global_log_table(), dur);
    p.durations[row_offset(rec.i2, (uint32_t)p.ld) + s] = dur[0];
}

__global__ void reset_counter_kernel(int32_t *counter) { *counter = 0; }

__global__ void philox_debug_kernel(int variant, int64_t n, const uint32_t *__restrict__ ctr,
                                    const uint32_t *__restrict__ key, uint32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (variant == 4) {
        const Words4 w = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i],
                                       key[2 * i + 1]);
        out[4 * i] = w.x;
        out[4 * i + 1] = w.y;
        out[4 * i + 2] = w.z;
        out[4 * i + 3] = w.w;
    } else {
        const Words2 w = philox2x32_10(ctr[2 * i], ctr[2 * i + 1], key[i]);
        out[2 * i] = w.x;
        out[2 * i + 1] = w.y;
    }
}

// The samplers' elementary functions on n inputs: 0 log_tab, 1 cos2pi_fast,
// 2 exp_fast.
__global__ void math_debug_kernel(int fn, int64_t n, const double *__restrict__ x, double *__restrict__ out) {
    __shared__ LogTableLds log_lds;
    stage_log_table(log_lds);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    out[i] = fn == 0 ? log_tab(v, log_lds.view()) : fn == 1 ? cos2pi_fast(v) : exp_fast(v);
}

__global__ void __launch_bounds__(256) copy_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                                   int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(double *dst, const double *src, int64_t n, hipStream_t stream) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    if (blocks > 0) hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, dst, src, n);
    return hipGetLastError();
}

template <bool kInjected, int kPaths, int kSpl, bool kPos>
const void *walk_symbol(int k) {
    switch (k) {
        case 4: return (const void *)walk_kernel<kInjected, kPaths, 4, kSpl, kPos>;
        case 8: return (const void *)walk_kernel<kInjected, kPaths, 8, kSpl, kPos>;
        case 12: return kSpl == 1 ? (const void *)walk_kernel<kInjected, kPaths, 12, 1, kPos> : nullptr;
        case 16: return kSpl == 1 ? (const void *)walk_kernel<kInjected, kPaths, 16, 1, kPos> : nullptr;
        default: return nullptr;
    }
}

template <bool kInjected, int kPaths, bool kPos>
const void *walk_symbol(int k, int spl) {
    return spl == 2   ? walk_symbol<kInjected, kPaths, 2, kPos>(k)
           : spl == 1 ? walk_symbol<kInjected, kPaths, 1, kPos>(k)
                      : nullptr;
}

// Winner positions (v.pos) only for sampled walks.
const void *walk_symbol(WalkVariant v) {
    if (v.injected)
        return v.pos ? nullptr
               : v.paths == kPathsFast ? walk_symbol<true, kPathsFast, false>(v.k, v.spl)
                                       : walk_symbol<true, kPathsGeneral, false>(v.k, v.spl);
    if (v.pos)
        return v.paths == kPathsFast ? walk_symbol<false, kPathsFast, true>(v.k, v.spl)
                                     : walk_symbol<false, kPathsGeneral, true>(v.k, v.spl);
    return v.paths == kPathsFast ? walk_symbol<false, kPathsFast, false>(v.k, v.spl)
                                 : walk_symbol<false, kPathsGeneral, false>(v.k, v.spl);
}

template <bool kInjected, int kPaths, int kSpl, bool kPos>
const void *walk_batch_symbol(int k) {
    switch (k) {
        case 4: return (const void *)walk_batch_kernel<kInjected, kPaths, 4, kSpl, kPos>;
        case 8: return (const void *)walk_batch_kernel<kInjected, kPaths, 8, kSpl, kPos>;
        case 12: return kSpl == 1 ? (const void *)walk_batch_kernel<kInjected, kPaths, 12, 1, kPos> : nullptr;
        case 16: return kSpl == 1 ? (const void *)walk_batch_kernel<kInjected, kPaths, 16, 1, kPos> : nullptr;
        default: return nullptr;
    }
}

template <bool kPos>
const void *walk_batch_symbol_of(WalkVariant v) {
    if (v.spl == 1)
        return v.paths == kPathsFast ? walk_batch_symbol<false, kPathsFast, 1, kPos>(v.k)
                                     : walk_batch_symbol<false, kPathsGeneral, 1, kPos>(v.k);
    if (v.spl == 2)
        return v.paths == kPathsFast ? walk_batch_symbol<false, kPathsFast, 2, kPos>(v.k)
                                     : walk_batch_symbol<false, kPathsGeneral, 2, kPos>(v.k);
    return nullptr;
}

// Batches run sampled plans (the injected variant has no batch form); v.pos:
// every plan's cause output is cause_pos.
const void *walk_batch_symbol(WalkVariant v) {
    if (v.injected) return nullptr;
    return v.pos ? walk_batch_symbol_of<true>(v) : walk_batch_symbol_of<false>(v);
}

hipError_t walk_batch_occupancy(WalkVariant v, int waves, size_t lds_bytes, int *blocks_per_cu) {
    const void *fn = walk_batch_symbol(v);
    if (!fn) return hipErrorInvalidValue;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, waves * kLanes, lds_bytes);
    if (e != hipSuccess) return e;
    hipFuncAttributes attr;
    e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) return e;
    if (attr.numRegs > 0) {
        const int granules = (attr.numRegs + 7) / 8;
        const int per_simd = std::min(8, 512 / (granules * 8));
        *blocks_per_cu = std::min(*blocks_per_cu, per_simd / ((waves + 3) / 4));
    }
    return hipSuccess;
}

hipError_t launch_walk_batch(const WalkBatchParams &b, WalkVariant v, int waves, int grid, size_t lds_bytes,
                             hipStream_t stream) {
    const void *fn = walk_batch_symbol(v);
    if (!fn) return hipErrorInvalidValue;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, b.counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    WalkBatchParams args = b;
    void *kargs[] = {&args};
    return hipLaunchKernel(fn, dim3(grid), dim3(waves * kLanes), kargs, lds_bytes, stream);
}

hipError_t launch_walk(const WalkParams &p, WalkVariant v, int grid, size_t lds_bytes, hipStream_t stream) {
    const void *fn = walk_symbol(v);
    if (!fn) return hipErrorInvalidValue;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    WalkParams args = p;
    void *kargs[] = {&args};
    return hipLaunchKernel(fn, dim3(grid), dim3(p.waves * kLanes), kargs, lds_bytes, stream);
}

// Workgroups of the walk kernel resident per CU.  The runtime's figure counts
// waves per CU; a workgroup's waves are spread over the CU's 4 SIMDs, ceil(W/4)
// on the fullest one, so a second workgroup fits only if every SIMD has that
// many wave slots left under the kernel's VGPR budget (512 per lane, in
// granules of 8, at most 8 waves per SIMD).
hipError_t walk_occupancy(WalkVariant v, int waves, size_t lds_bytes, int *blocks_per_cu) {
    const void *fn = walk_symbol(v);
    if (!fn) return hipErrorInvalidValue;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, waves * kLanes, lds_bytes);
    if (e != hipSuccess) return e;
    hipFuncAttributes attr;
    e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) return e;
    if (attr.numRegs > 0) {
        const int granules = (attr.numRegs + 7) / 8;
        const int per_simd = std::min(8, 512 / (granules * 8));
        const int per_wg = (waves + 3) / 4;
        *blocks_per_cu = std::min(*blocks_per_cu, per_simd / per_wg);
    }
    return hipSuccess;
}

hipError_t launch_walk_wide(const WalkParams &p, bool injected, int32_t e0, int32_t n_events, int32_t groups,
                            hipStream_t stream) {
    if (n_events <= 0 || groups <= 0) return hipSuccess;
    if (injected)
        hipLaunchKernelGGL(walk_wide_kernel<true>, dim3((unsigned)n_events, (unsigned)groups), dim3(kLanes), 0, stream,
                           p, e0);
    else
        hipLaunchKernelGGL(walk_wide_kernel<false>, dim3((unsigned)n_events, (unsigned)groups), dim3(kLanes), 0, stream,
                           p, e0);
    return hipGetLastError();
}

hipError_t launch_walk_wide_orphans(const WalkParams &p, int32_t groups, hipStream_t stream) {
    if (p.n_orphans <= 0 || groups <= 0 || !p.durations || !p.seeds) return hipSuccess;
    hipLaunchKernelGGL(walk_wide_orphans_kernel, dim3((unsigned)p.n_orphans, (unsigned)groups), dim3(kLanes), 0, stream,
                       p);
    return hipGetLastError();
}

hipError_t launch_philox_debug(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out,
                               hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    if (grid > 0) hipLaunchKernelGGL(philox_debug_kernel, dim3(grid), dim3(block), 0, stream, variant, n, ctr, key, out);
    return hipGetLastError();
}

hipError_t launch_math_debug(int fn, int64_t n, const double *x, double *out, hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    if (grid > 0) hipLaunchKernelGGL(math_debug_kernel, dim3(grid), dim3(block), 0, stream, fn, n, x, out);
    return hipGetLastError();
}

}  // namespace mcdp
