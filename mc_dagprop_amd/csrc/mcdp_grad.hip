// gfx950 kernels of the reverse (pathwise-gradient) walk.
//
// reverse_kernel: one workgroup per tile of 64 samples (lane l carries sample
// tile * 64 + l), W waves splitting each step's events as in the forward walk,
// the steps in reverse topological order.  Per event e and sample:
//   g      = sum over e's successor records (t, k) of adj[t] if win[t] == k
//   active = realized[e] > earliest[e]
//   gt     = g + (active ? w[e] : 0)                   (dJ / d realized[e])
//   J     += active ? w[e] (realized[e] - earliest[e]) : 0
//   dE[e] += (active ? -w[e] : 0) + (capped or cause -1 ? gt : 0)
//   winner: the last entry k whose source is cause[e] (and, for a source
//           listed more than once, whose capped candidate equals realized[e])
//   entry k == winner: dD[link_k] += gt * d duration_k / d d; a stale entry
//           also gives dE[src_k] += gt, any other passes gt to src_k's adjoint
//   adj[e] = gt, win[e] = winner (-1 if none, capped or stale)  -> e's slot
// Sums over the tile's lanes are wave reductions (DPP, a fixed order), written
// once per (tile, column) to the partial matrix; reduce_columns sums the tiles.
// The arithmetic is FP64 add, multiply and compare-select; this file is
// compiled without FMA contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mcdp_grad.hpp"
#include "mcdp_kernels.hpp"
#include "mcdp_sampler.hpp"

namespace mcdp {

namespace {

__device__ __forceinline__ uint64_t row_offset(int32_t row, uint32_t ld) { return (uint64_t)(uint32_t)row * ld; }

// One DPP move of a double (both halves) with the given control; lanes the
// row mask leaves out read 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, kCtrl, kRowMask, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), kCtrl, kRowMask, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Sum of v over the 64 lanes, in a fixed order, as a wave-uniform value:
// pairs and quads (quad_perm), halves of a row (row_half_mirror), rows of 16
// (row_mirror), then row 1 += row 0 and row 3 += row 2 (row_bcast:15) and
// row 3 += rows 0-1 (row_bcast:31); lane 63 holds the total.
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_f64<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
    v = v + dpp_f64<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
    v = v + dpp_f64<0x141, 0xF>(v);  // row_half_mirror
    v = v + dpp_f64<0x140, 0xF>(v);  // row_mirror
    v = v + dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = v + dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ bool any_lane(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }

// A slot's adjoint and winner position for this lane: LDS for the first
// lds_slots slots, the workgroup's global region for the rest.
struct Slots {
    double *g_lds;
    int32_t *w_lds;
    double *g_glb;
    int32_t *w_glb;
    int32_t n_lds;
    int lane;
    __device__ __forceinline__ void read(int32_t slot, double &g, int32_t &w) const {
        if (slot < n_lds) {
            g = g_lds[slot * kLanes + lane];
            w = w_lds[slot * kLanes + lane];
        } else {
            const int32_t s = slot - n_lds;
            g = g_glb[s * kLanes + lane];
            w = w_glb[s * kLanes + lane];
        }
    }
    __device__ __forceinline__ void write(int32_t slot, double g, int32_t w) const {
        if (slot < n_lds) {
            g_lds[slot * kLanes + lane] = g;
            w_lds[slot * kLanes + lane] = w;
        } else {
            const int32_t s = slot - n_lds;
            g_glb[s * kLanes + lane] = g;
            w_glb[s * kLanes + lane] = w;
        }
    }
};

__global__ void __launch_bounds__(kMaxWaves * kLanes) reverse_kernel(RevParams p) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = p.waves;
    const uint32_t ld = (uint32_t)p.ld;
    const auto step_off = as_const_space(p.step_off);
    const auto recs = as_const_space(p.recs);
    const auto ev_off = as_const_space(p.ev_off);
    const auto ev_list = as_const_space(p.ev_list);
    const auto ev_slot = as_const_space(p.ev_slot);
    const auto weights = as_const_space(p.weights);
    const auto earliest = as_const_space(p.earliest);

    Slots sl;
    sl.g_lds = lds;
    sl.w_lds = reinterpret_cast<int32_t *>(lds + (size_t)p.lds_slots * kLanes);
    sl.g_glb = p.gslots + (size_t)blockIdx.x * p.global_slots * kLanes * 2;
    sl.w_glb = reinterpret_cast<int32_t *>(sl.g_glb + (size_t)p.global_slots * kLanes);
    sl.n_lds = p.lds_slots;
    sl.lane = lane;

    for (int32_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x) {
        // Lanes past the last sample repeat sample n - 1 with weight 0: every
        // sum they enter gets +0 from them.
        const int32_t s_raw = (p.tile0 + tile) * kLanes + lane;
        const bool live = s_raw < p.n;
        const uint32_t s = live ? (uint32_t)s_raw : (uint32_t)(p.n - 1);
        double *const prow = p.partial + (size_t)tile * p.n_cols;
        double obj = 0.0;

        for (int t = 0; t < p.n_steps; ++t) {
            const int e0 = ev_off[t * W + wave];
            const int ne = ev_off[t * W + wave + 1] - e0;
            // This wave's events of the step: realized and cause loads issued together.
            double rr[kRevEventsPerWave];
            int32_t cc[kRevEventsPerWave];
#pragma unroll
            for (int j = 0; j < kRevEventsPerWave; ++j) {
                rr[j] = 0.0;
                cc[j] = -1;
                if (j < ne) {
                    const uint64_t row = row_offset(ev_list[e0 + j], ld);
                    rr[j] = p.realized[row + s];
                    cc[j] = p.cause[row + s];
                }
            }
            double gres[kRevEventsPerWave];
            int32_t wres[kRevEventsPerWave];
            int h = step_off[t * W + wave];
#pragma unroll
            for (int j = 0; j < kRevEventsPerWave; ++j) {
                gres[j] = 0.0;
                wres[j] = -1;
                if (j >= ne) continue;
                const Record hdr = recs[h];
                const int32_t ev = hdr.i0;
                const int32_t n_succ = hdr.i2;
                const int32_t n_pred = (int32_t)(hdr.info & kRevNpredMask);
                const double we = weights[ev];
                // adjoint from the successors
                double g = 0.0;
                for (int q = h + 1; q <= h + n_succ; ++q) {
                    const Record sr = recs[q];
                    double gv;
                    int32_t wv;
                    sl.read(sr.i1, gv, wv);
                    g += (wv == sr.i2) ? gv : 0.0;
                }
                const double r = rr[j];
                const int32_t c = cc[j];
                const bool active = live && (r > hdr.a);
                const double gt = g + (active ? we : 0.0);
                obj += active ? we * (r - hdr.a) : 0.0;
                const bool capped = (c < 0) || (r == hdr.b);
                double ge = active ? -we : 0.0;
                ge += capped ? gt : 0.0;
                // the winning entry
                const int pb = h + 1 + n_succ;
                int32_t win = -1;
                if (any_lane(live && !capped)) {
                    for (int k = 0; k < n_pred; ++k) {
                        const Record pr = recs[pb + k];
                        bool m = live && !capped && (c == pr.i0);
                        if ((pr.info & kRevDup) && any_lane(m)) {
                            // a repeated source: the entry whose candidate is the maximum
                            const double base = (pr.info & kRevStale) ? earliest[pr.i0]
                                                                      : p.realized[row_offset(pr.i0, ld) + s];
                            const double d = (pr.info & kInfoEdgeValid) ? p.durations[row_offset(pr.i2, ld) + s] : 0.0;
                            double tt = base + d;
                            tt = (hdr.b < tt) ? hdr.b : tt;
                            m = m && (tt == r);
                        }
                        win = m ? k : win;
                    }
                }
                // derivatives of the entries; the winner passes the adjoint on
                int32_t wslot = -1;
                for (int k = 0; k < n_pred; ++k) {
                    const Record pr = recs[pb + k];
                    const bool mine = win == k;
                    const bool stale = (pr.info & kRevStale) != 0;
                    if (mine && !stale) wslot = k;
                    const uint32_t cls = (pr.info & kRevClassMask) >> kRevClassShift;
                    if (cls == kMulNone && !stale) continue;
                    double dm = 0.0, ds = 0.0;
                    if (any_lane(mine)) {
                        if (cls != kMulNone) {
                            double mult = 1.0;
                            if (cls == kMulRead) {
                                mult = p.durations[row_offset(pr.i2, ld) + s] * pr.a;
                            } else if (cls == kMulDraw) {
                                const uint32_t seed[1] = {(uint32_t)p.seeds[s]};
                                double unit[1];
                                draw_durations<1>(pr.info & ~(kRevClassMask | kRevDup | kRevStale), 1.0, pr.b,
                                                  (uint32_t)pr.i2, seed, as_const_space(p.alias),
                                                  as_const_space(p.dists), unit);
                                mult = unit[0];
                            }
                            dm = wave_sum(mine ? gt * mult : 0.0);
                        }
                        if (stale) ds = wave_sum(mine ? gt : 0.0);
                    }
                    if (lane == 0) {
                        prow[pr.i1] = dm;
                        if (stale) prow[pr.i1 + 1] = ds;
                    }
                }
                const double se = any_lane(ge != 0.0) ? wave_sum(ge) : 0.0;
                if (lane == 0) prow[p.L + ev] = se;
                gres[j] = gt;
                wres[j] = wslot;
                h = pb + n_pred;
            }
            // Every slot read of this step is done before any is overwritten,
            // and every commit is visible before the next step reads it.
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kRevEventsPerWave; ++j) {
                if (j < ne) {
                    const int32_t slot = ev_slot[e0 + j];
                    if (slot >= 0) sl.write(slot, gres[j], wres[j]);
                }
            }
            __syncthreads();
        }
        const double so = wave_sum(obj);
        if (lane == 0) prow[p.obj_col + wave] = so;
    }
}

__global__ void __launch_bounds__(256) reduce_columns_kernel(const double *__restrict__ partial, int32_t n_tiles,
                                                             int32_t n_cols, const int32_t *__restrict__ map_off,
                                                             const int32_t *__restrict__ map_col, int32_t L,
                                                             int32_t E, int accumulate, double *grad_dmin,
                                                             double *grad_earliest, double *objective) {
    const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o > L + E) return;
    double *const out = o < L ? grad_dmin + o : (o < L + E ? grad_earliest + (o - L) : objective);
    const int32_t c0 = map_off[o], c1 = map_off[o + 1];
    double acc = accumulate ? *out : 0.0;
    for (int32_t t = 0; t < n_tiles; ++t) {
        const double *row = partial + (size_t)t * n_cols;
        for (int32_t c = c0; c < c1; ++c) acc += row[map_col[c]];
    }
    *out = acc;
}

}  // namespace

hipError_t launch_reverse(const RevParams &p, int grid, size_t lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL(reverse_kernel, dim3((unsigned)grid), dim3((unsigned)(p.waves * kLanes)), lds_bytes, stream, p);
    return hipGetLastError();
}

hipError_t reverse_occupancy(int waves, size_t lds_bytes, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, reverse_kernel, waves * kLanes, lds_bytes);
}

hipError_t launch_reduce_columns(const double *partial, int32_t n_tiles, int32_t n_cols, const int32_t *map_off,
                                 const int32_t *map_col, int32_t L, int32_t E, int accumulate, double *grad_dmin,
                                 double *grad_earliest, double *objective, hipStream_t stream) {
    const int32_t n_out = L + E + 1;
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, stream, partial,
                       n_tiles, n_cols, map_off, map_col, L, E, accumulate, grad_dmin, grad_earliest, objective);
    return hipGetLastError();
}

}  // namespace mcdp
