// gfx950 kernels of the reverse (pathwise-gradient) walk.
//
// reverse_kernel: one workgroup per tile of 64 spl samples (lane l carries
// samples tile * 64 spl + i * 64 + l, i < spl), W waves splitting each step's
// events as in the forward walk,
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
// An event's per-sample rows (realized, cause, durations) are loaded while the
// previous event is processed.
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

// Issue priority from the events a wave has left in its step (5 or more: 3,
// 3-4: 2, 2: 1, the last: 0).  A SIMD otherwise issues its oldest waves
// first, so with equal work its youngest wave finishes every step last while
// the others wait at the step's barrier; favouring the waves furthest behind
// lets a SIMD's waves reach the barrier together.
__device__ __forceinline__ void set_priority(int left) {
    if (left >= 5)
        __builtin_amdgcn_s_setprio(3);
    else if (left >= 3)
        __builtin_amdgcn_s_setprio(2);
    else if (left >= 2)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

__device__ __forceinline__ bool any_lane(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }

// A slot's adjoints and winner positions for this lane's kSpl samples
// (adjacent: one ds_read_b64 / b128 per plane).  kGlobal: slots past n_lds
// live in the workgroup's global region; otherwise every slot is in LDS.
template <bool kGlobal, int kSpl>
struct Slots {
    double *g_lds;
    int32_t *w_lds;
    double *g_glb;
    int32_t *w_glb;
    int32_t n_lds;
    int lane;
    __device__ __forceinline__ void read(int32_t slot, double (&g)[kSpl], int32_t (&w)[kSpl]) const {
        const bool in_lds = !kGlobal || slot < n_lds;
        const int32_t i = ((in_lds ? slot : slot - n_lds) * kLanes + lane) * kSpl;
        if (in_lds) {
            if constexpr (kSpl == 1) {
                g[0] = g_lds[i];
                w[0] = w_lds[i];
            } else {
                const double2 gg = *reinterpret_cast<const double2 *>(g_lds + i);
                const int2 ww = *reinterpret_cast<const int2 *>(w_lds + i);
                g[0] = gg.x, g[1] = gg.y, w[0] = ww.x, w[1] = ww.y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kSpl; ++k) {
                g[k] = __builtin_nontemporal_load(g_glb + i + k);
                w[k] = __builtin_nontemporal_load(w_glb + i + k);
            }
        }
    }
    __device__ __forceinline__ void write(int32_t slot, const double (&g)[kSpl], const int32_t (&w)[kSpl]) const {
        const bool in_lds = !kGlobal || slot < n_lds;
        const int32_t i = ((in_lds ? slot : slot - n_lds) * kLanes + lane) * kSpl;
        if (in_lds) {
            if constexpr (kSpl == 1) {
                g_lds[i] = g[0];
                w_lds[i] = w[0];
            } else {
                *reinterpret_cast<double2 *>(g_lds + i) = double2{g[0], g[1]};
                *reinterpret_cast<int2 *>(w_lds + i) = int2{w[0], w[1]};
            }
        } else {
#pragma unroll
            for (int k = 0; k < kSpl; ++k) {
                g_glb[i + k] = g[k];
                w_glb[i + k] = w[k];
            }
        }
    }
};

// Column sums of one wave (wave-uniform values), staged one per lane and
// written with one store instruction per 64 of them after the step's loads:
// a store issued between loads would make the next load's wait include it.
struct ColumnStage {
    int32_t col = 0;
    double val = 0.0;
    int n = 0;  // wave-uniform
    __device__ __forceinline__ void flush(double *prow, int lane) {
        if (lane < n) prow[col] = val;
        n = 0;
    }
    __device__ __forceinline__ void put(int32_t c, double v, double *prow, int lane) {
        const bool mine = lane == n;
        col = mine ? c : col;
        val = mine ? v : val;
        if (++n == kLanes) flush(prow, lane);
    }
};

// Eight 32-bit values with a wave-uniform dynamic index, kept in registers:
// the index selects a branch, not an address, so no private array is needed.
struct Regs8 {
    int32_t v0 = -1, v1 = -1, v2 = -1, v3 = -1, v4 = -1, v5 = -1, v6 = -1, v7 = -1;
    __device__ __forceinline__ void set(int j, int32_t x) {
        switch (j) {
            case 0: v0 = x; break;
            case 1: v1 = x; break;
            case 2: v2 = x; break;
            case 3: v3 = x; break;
            case 4: v4 = x; break;
            case 5: v5 = x; break;
            case 6: v6 = x; break;
            default: v7 = x; break;
        }
    }
    __device__ __forceinline__ int32_t get(int j) const {
        switch (j) {
            case 0: return v0;
            case 1: return v1;
            case 2: return v2;
            case 3: return v3;
            case 4: return v4;
            case 5: return v5;
            case 6: return v6;
            default: return v7;
        }
    }
};

// The slots of a (step, wave)'s events, read as one row.
template <int K>
struct EventSlots {
    int32_t v[K];
};

constexpr int kPre = 3;        // multiplier rows of an event's first predecessors, loaded one event ahead
constexpr int kFastPreds = 4;  // predecessor records read per group

// Lane sum of a per-sample value, then the wave sum: a fixed order.
template <int kSpl>
__device__ __forceinline__ double tile_sum(const double (&v)[kSpl]) {
    double x = v[0];
#pragma unroll
    for (int i = 1; i < kSpl; ++i) x += v[i];
    return wave_sum(x);
}

template <int kSpl>
__device__ __forceinline__ bool any_of(const bool (&b)[kSpl]) {
    bool x = b[0];
#pragma unroll
    for (int i = 1; i < kSpl; ++i) x = x || b[i];
    return any_lane(x);
}

template <bool kGlobal, int kSpl>
__global__ void __launch_bounds__(kMaxWaves * kLanes) reverse_kernel(RevParams p) {
    extern __shared__ double lds[];
    constexpr int K = kRevEventsPerWave;
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = p.waves;
    const uint32_t ld = (uint32_t)p.ld;
    const auto blk_off = as_const_space(p.blk_off);
    const auto ev_off = as_const_space(p.ev_off);
    const auto col_off = as_const_space(p.col_off);
    const auto weights = as_const_space(p.weights);
    const auto earliest = as_const_space(p.earliest);
    const auto blocks_s = as_const_space(p.blocks);

    Slots<kGlobal, kSpl> sl;
    sl.g_lds = lds;
    sl.w_lds = reinterpret_cast<int32_t *>(lds + (size_t)p.lds_slots * kLanes * kSpl);
    sl.g_glb = p.gslots + (size_t)blockIdx.x * p.global_slots * kLanes * kSpl * 2;
    sl.w_glb = reinterpret_cast<int32_t *>(sl.g_glb + (size_t)p.global_slots * kLanes * kSpl);
    sl.n_lds = p.lds_slots;
    sl.lane = lane;

    // A block's first kBlkWindow dwords, one per lane (vector load), and its
    // fields by lane index; dwords past the window come through scalar loads.
    auto window = [&](int32_t pos) -> uint32_t { return p.blocks[pos + lane]; };
    auto dw = [&](uint32_t win, int32_t pos, int k) -> uint32_t {
        return k < kBlkWindow ? (uint32_t)__builtin_amdgcn_readlane((int)win, k) : blocks_s[pos + k];
    };
    auto f64 = [&](uint32_t win, int32_t pos, int k) -> double {
        const uint64_t lo = dw(win, pos, k), hi = dw(win, pos, k + 1);
        return __longlong_as_double((long long)((hi << 32) | lo));
    };

    for (int32_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x) {
        // Lane l carries samples (tile0 + tile) * 64 kSpl + i * 64 + l.  Lanes
        // past the last sample repeat sample n - 1 with weight 0: every sum
        // they enter gets +0 from them.
        bool live[kSpl];
        uint32_t s[kSpl], s_unit[kSpl];
#pragma unroll
        for (int i = 0; i < kSpl; ++i) {
            const int32_t s_raw = (p.tile0 + tile) * (kLanes * kSpl) + i * kLanes + lane;
            live[i] = s_raw < p.n;
            s[i] = live[i] ? (uint32_t)s_raw : (uint32_t)(p.n - 1);
            s_unit[i] = s[i] - (uint32_t)p.tile0 * (kLanes * kSpl);
        }
        double *const prow = p.partial + (size_t)tile * p.n_cols;
        double obj[kSpl];
#pragma unroll
        for (int i = 0; i < kSpl; ++i) obj[i] = 0.0;
        ColumnStage st;
        auto mult_row = [&](uint32_t info, uint32_t edge, double (&out)[kSpl]) {
            const bool unit = (info & kRevClassMask) == (kMulDraw << kRevClassShift);
            const double *base = unit ? p.unit + row_offset((int32_t)((info >> kInfoDistShift) & kInfoDistMask), p.ld_unit)
                                      : p.durations + row_offset((int32_t)edge, ld);
#pragma unroll
            for (int i = 0; i < kSpl; ++i) out[i] = base[unit ? s_unit[i] : s[i]];
        };
        auto reads_row = [](uint32_t info) {
            const uint32_t cls = (info & kRevClassMask) >> kRevClassShift;
            return cls == kMulRead || cls == kMulDraw;
        };

        // The event pipeline over a wave's stream of the whole tile: blocks two
        // events ahead, per-sample rows (realized, cause, the multiplier rows
        // of the first kPre predecessors) and the weight one event ahead.
        struct Rows {
            double r[kSpl];
            int32_t c[kSpl];
            double d[kPre][kSpl];
        };
        // first block and event count of step tt in this wave's stream
        auto first_of = [&](int tt, int &ne_out) -> int32_t {
            ne_out = ev_off[tt * W + wave + 1] - ev_off[tt * W + wave];
            return blk_off[tt * W + wave];
        };
        // A cursor walks the wave's events in stream order: step t, event j of
        // the step's ne, block at dword pos.
        struct Cursor {
            int t, j, ne;
            int32_t pos;
        };
        auto advance = [&](Cursor &c, uint32_t size) {
            // c is at an event whose block has `size` dwords
            ++c.j;
            c.pos += (int32_t)size;
            while (c.t < p.n_steps && c.j >= c.ne) {
                ++c.t;
                c.j = 0;
                if (c.t < p.n_steps) c.pos = first_of(c.t, c.ne);
            }
        };
        Cursor c0{0, 0, 0, 0};
        if (p.n_steps > 0) c0.pos = first_of(0, c0.ne);
        while (c0.t < p.n_steps && c0.ne == 0) {
            ++c0.t;
            if (c0.t < p.n_steps) c0.pos = first_of(c0.t, c0.ne);
        }
        const bool any_event = c0.t < p.n_steps;
        // blocks of the current event and the next one, rows of the current event
        uint32_t win0 = any_event ? window(c0.pos) : 0u;
        Cursor c1 = c0;
        uint32_t win1 = 0;
        bool have1 = false;
        if (any_event) {
            advance(c1, dw(win0, c0.pos, 3));
            have1 = c1.t < p.n_steps;
            if (have1) win1 = window(c1.pos);
        }
        // A fixed number of loads per event (an entry without a multiplier row
        // reads the event's realized row instead, and its value is never used),
        // so the waits for one event's rows can leave the next event's loads in
        // flight.
        auto load_rows = [&](uint32_t win, int32_t pos, Rows &x) {
            const uint64_t row = row_offset((int32_t)dw(win, pos, 0), ld);
#pragma unroll
            for (int i = 0; i < kSpl; ++i) {
                x.r[i] = p.realized[row + s[i]];
                x.c[i] = p.cause[row + s[i]];
            }
            const int n_pred = (int)(dw(win, pos, 2) & 0xFFFFu);
#pragma unroll
            for (int k = 0; k < kPre; ++k) {
                const uint32_t info = k < n_pred ? dw(win, pos, kBlkHeader + kBlkPred * k + 4) : 0u;
                const uint32_t edge = k < n_pred ? dw(win, pos, kBlkHeader + kBlkPred * k + 3) : 0u;
                const bool unit = (info & kRevClassMask) == (kMulDraw << kRevClassShift);
                const bool rd = k < n_pred && reads_row(info);
                const double *base = !rd   ? p.realized + row
                                     : unit ? p.unit + row_offset((int32_t)((info >> kInfoDistShift) & kInfoDistMask),
                                                                  p.ld_unit)
                                            : p.durations + row_offset((int32_t)edge, ld);
#pragma unroll
                for (int i = 0; i < kSpl; ++i) x.d[k][i] = base[(rd && unit) ? s_unit[i] : s[i]];
            }
        };
        Rows cur{}, nxt{};
        double we_cur = 0.0, we_nxt = 0.0;
        if (any_event) {
            load_rows(win0, c0.pos, cur);
            we_cur = weights[(int32_t)dw(win0, c0.pos, 0)];
        }

        // A step's event range and first column, one per lane (lanes 0-2), loaded a
        // step ahead with one vector load.
        auto step_meta = [&](int tt) -> int32_t {
            const int i = (tt < p.n_steps ? tt : 0) * W + wave;
            return lane < 2 ? p.ev_off[i + lane] : p.col_off[i];
        };
        int32_t meta = p.n_steps > 0 ? step_meta(0) : 0;
        for (int ts = 0; ts < p.n_steps; ++ts) {
            const int e0 = __builtin_amdgcn_readlane(meta, 0);
            const int ne = __builtin_amdgcn_readlane(meta, 1) - e0;
            int32_t ev_col = __builtin_amdgcn_readlane(meta, 2);  // the next column this wave writes
            meta = step_meta(ts + 1);
            static_assert(K == 8, "Regs8 holds one step's winner positions");
            double gres[K][kSpl];
            Regs8 wres;
#pragma unroll 1
            for (int jj = 0; jj < ne; ++jj) {
                set_priority(ne - jj);
                // c0 / win0 / cur: this event; c1 / win1: the next; start the one after and the next rows
                const int32_t pos = c0.pos;
                const uint32_t win = win0;
                Cursor c2 = c1;
                uint32_t win2 = 0;
                bool have2 = false;
                if (have1) {
                    advance(c2, dw(win1, c1.pos, 3));
                    have2 = c2.t < p.n_steps;
                }
                // the same loads whether or not a further event exists (block 0 stands in)
                win2 = window(have2 ? c2.pos : 0);
                load_rows(win1, have1 ? c1.pos : 0, nxt);
                we_nxt = weights[(int32_t)dw(win1, have1 ? c1.pos : 0, 0)];

                const uint32_t counts = dw(win, pos, 2);
                const int32_t n_pred = (int32_t)(counts & 0xFFFFu);
                const int32_t n_succ = (int32_t)(counts >> 16);
                const double ea = f64(win, pos, 4), eb = f64(win, pos, 6);
                const double we = we_cur;
                const int sb = kBlkHeader + kBlkPred * n_pred;  // first successor dword
                // adjoint from the successors, four entries per round
                double g[kSpl];
#pragma unroll
                for (int i = 0; i < kSpl; ++i) g[i] = 0.0;
                {
                    int q = 0;
                    for (; q + 4 <= n_succ; q += 4) {
                        int32_t sx[4], sy[4];
                        double gv[4][kSpl];
                        int32_t wv[4][kSpl];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            sx[u] = (int32_t)dw(win, pos, sb + kBlkSucc * (q + u));
                            sy[u] = (int32_t)dw(win, pos, sb + kBlkSucc * (q + u) + 1);
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) sl.read(sx[u], gv[u], wv[u]);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int i = 0; i < kSpl; ++i) g[i] += (wv[u][i] == sy[u]) ? gv[u][i] : 0.0;
                    }
                    for (; q < n_succ; ++q) {
                        const int32_t sx = (int32_t)dw(win, pos, sb + kBlkSucc * q);
                        const int32_t sy = (int32_t)dw(win, pos, sb + kBlkSucc * q + 1);
                        double gv[kSpl];
                        int32_t wv[kSpl];
                        sl.read(sx, gv, wv);
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) g[i] += (wv[i] == sy) ? gv[i] : 0.0;
                    }
                }
                double gt[kSpl], ge[kSpl];
                bool capped[kSpl], open[kSpl], ge_nz[kSpl];
#pragma unroll
                for (int i = 0; i < kSpl; ++i) {
                    const double r = cur.r[i];
                    const bool active = live[i] && (r > ea);
                    capped[i] = (cur.c[i] < 0) || (r == eb);
                    open[i] = live[i] && !capped[i];
                    gt[i] = g[i] + (active ? we : 0.0);
                    obj[i] += active ? we * (r - ea) : 0.0;
                    ge[i] = active ? -we : 0.0;
                    ge[i] += capped[i] ? gt[i] : 0.0;
                    ge_nz[i] = ge[i] != 0.0;
                }
                const bool any_open = any_of<kSpl>(open);

                // winners: the last entry whose source is the cause (for a repeated
                // source, whose capped candidate is the maximum)
                int32_t win_k[kSpl];
#pragma unroll
                for (int i = 0; i < kSpl; ++i) win_k[i] = -1;
                if (any_open) {
                    for (int k = 0; k < n_pred; ++k) {
                        const int kb = kBlkHeader + kBlkPred * k;
                        const int32_t src = (int32_t)dw(win, pos, kb + 2);
                        const uint32_t info = dw(win, pos, kb + 4);
                        bool m[kSpl];
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) m[i] = open[i] && (cur.c[i] == src);
                        if ((info & kRevDup) && any_of<kSpl>(m)) {
                            const uint32_t edge = dw(win, pos, kb + 3);
#pragma unroll
                            for (int i = 0; i < kSpl; ++i) {
                                const double base =
                                    (info & kRevStale) ? earliest[src] : p.realized[row_offset(src, ld) + s[i]];
                                const double d =
                                    (info & kInfoEdgeValid) ? p.durations[row_offset((int32_t)edge, ld) + s[i]] : 0.0;
                                double tt = base + d;
                                tt = (eb < tt) ? eb : tt;
                                m[i] = m[i] && (tt == cur.r[i]);
                            }
                        }
#pragma unroll
                        for (int i = 0; i < kSpl; ++i) win_k[i] = m[i] ? k : win_k[i];
                    }
                }
                // the derivative columns of each entry; the winner passes the adjoint on
                int32_t wslot[kSpl];
#pragma unroll
                for (int i = 0; i < kSpl; ++i) wslot[i] = -1;
                for (int k = 0; k < n_pred; ++k) {
                    const int kb = kBlkHeader + kBlkPred * k;
                    const uint32_t info = dw(win, pos, kb + 4);
                    const bool stale = (info & kRevStale) != 0;
                    bool mine[kSpl];
#pragma unroll
                    for (int i = 0; i < kSpl; ++i) {
                        mine[i] = win_k[i] == k;
                        if (mine[i] && !stale) wslot[i] = k;
                    }
                    const uint32_t cls = (info & kRevClassMask) >> kRevClassShift;
                    if (cls == kMulNone && !stale) continue;
                    double dm = 0.0, ds = 0.0;
                    if (any_of<kSpl>(mine)) {
                        if (cls != kMulNone) {
                            double dk[kSpl];
                            if (k < kPre) {
#pragma unroll
                                for (int i = 0; i < kSpl; ++i) dk[i] = cur.d[k < kPre ? k : 0][i];
                            } else if (reads_row(info)) {
                                mult_row(info, dw(win, pos, kb + 3), dk);
                            } else {
#pragma unroll
                                for (int i = 0; i < kSpl; ++i) dk[i] = 0.0;
                            }
                            const double a = f64(win, pos, kb);
                            double v[kSpl];
#pragma unroll
                            for (int i = 0; i < kSpl; ++i) v[i] = mine[i] ? gt[i] * (cls == kMulOne ? 1.0 : dk[i] * a) : 0.0;
                            dm = tile_sum<kSpl>(v);
                        }
                        if (stale) {
                            double v[kSpl];
#pragma unroll
                            for (int i = 0; i < kSpl; ++i) v[i] = mine[i] ? gt[i] : 0.0;
                            ds = tile_sum<kSpl>(v);
                        }
                    }
                    st.put(ev_col, dm, prow, lane);
                    if (stale) st.put(ev_col + 1, ds, prow, lane);
                    ev_col += stale ? 2 : 1;
                }
                // the event's earliest-time column follows its entries' columns
                st.put(ev_col, any_of<kSpl>(ge_nz) ? tile_sum<kSpl>(ge) : 0.0, prow, lane);
                ++ev_col;
#pragma unroll
                for (int i = 0; i < kSpl; ++i) gres[jj][i] = gt[i];
                wres.set(jj, kSpl == 1 ? wslot[0]
                                       : (int32_t)(((uint32_t)wslot[0] & 0xFFFFu) | ((uint32_t)wslot[kSpl - 1] << 16)));
                // shift the pipeline
                c0 = c1;
                win0 = win1;
                c1 = c2;
                win1 = win2;
                have1 = have2;
                cur = nxt;
                we_cur = we_nxt;
            }
            st.flush(prow, lane);
            // The step's slots in one scalar load (the list ends in K padding
            // entries), in flight across the barrier: no wait on the stores.
            const EventSlots<K> srow = *as_const_space(reinterpret_cast<const EventSlots<K> *>(p.ev_slot + e0));
            // Every slot read of this step is done before any is overwritten,
            // and every commit is visible before the next step reads it.
            __syncthreads();
#pragma unroll
            for (int jj = 0; jj < K; ++jj) {
                const int32_t slot = srow.v[jj];
                if (jj < ne && slot >= 0) {
                    int32_t w[kSpl];
                    const int32_t x = wres.get(jj);
                    w[0] = kSpl == 1 ? x : (int32_t)(int16_t)(x & 0xFFFF);
                    if constexpr (kSpl == 2) w[1] = x >> 16;
                    sl.write(slot, gres[jj], w);
                }
            }
            __syncthreads();
        }
        st.put(p.obj_col + wave, tile_sum<kSpl>(obj), prow, lane);
        st.flush(prow, lane);
    }
}

// Unit durations of the links that need them (relative kinds with minimal
// duration 0): out[q][i] = the duration link q's draw gives sample
// first + i (clamped to n - 1) at minimal duration 1, i.e. 1 + X rounded once.
__global__ void __launch_bounds__(256) unit_draw_kernel(const Record *__restrict__ links, int32_t n_links,
                                                        const int32_t *__restrict__ seeds, int32_t first, int32_t count,
                                                        int32_t n, const DistParams *dists, const AliasRec *alias,
                                                        double *__restrict__ out, uint32_t ld_out) {
    __shared__ LogTableLds log_lds;
    stage_log_table(log_lds);
    __syncthreads();
    const int32_t q = blockIdx.y;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_links || i >= count) return;
    const int32_t si = first + i < n ? first + i : n - 1;
    const Record rec = as_const_space(links)[q];
    const uint32_t seed[1] = {(uint32_t)seeds[si]};
    double unit[1];
    draw_durations<1>(rec.info, 1.0, rec.b, (uint32_t)rec.i2, seed, as_const_space(alias), as_const_space(dists),
                      log_lds.view(), unit);
    out[row_offset(q, ld_out) + i] = unit[0];
}

// groups[g][c] = sum over tiles t of group g (kReduceGroup consecutive tiles),
// in tile order, of partial[t][c]: the first of the two reduction stages,
// with a group's loads in flight together.
__global__ void __launch_bounds__(256) group_sum_kernel(const double *__restrict__ partial, int32_t n_tiles,
                                                        int32_t n_cols, double *__restrict__ groups) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t g = blockIdx.y;
    if (c >= n_cols) return;
    const int32_t t0 = g * kReduceGroup;
    const int32_t m = min(kReduceGroup, n_tiles - t0);
    const double *col = partial + (size_t)t0 * n_cols + c;
    double v[kReduceGroup];
#pragma unroll
    for (int i = 0; i < kReduceGroup; ++i) v[i] = i < m ? col[(size_t)i * n_cols] : 0.0;
    double acc = v[0];
#pragma unroll
    for (int i = 1; i < kReduceGroup; ++i)
        if (i < m) acc += v[i];
    groups[(size_t)g * n_cols + c] = acc;
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
    if (c1 == c0 + 1) {  // one column (most outputs): 16 tiles' loads in flight, added in tile order
        const double *col = partial + map_col[c0];
        int32_t t = 0;
        for (; t + 16 <= n_tiles; t += 16) {
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = col[(size_t)(t + i) * n_cols];
#pragma unroll
            for (int i = 0; i < 16; ++i) acc += v[i];
        }
        for (; t < n_tiles; ++t) acc += col[(size_t)t * n_cols];
    } else {
        for (int32_t t = 0; t < n_tiles; ++t) {
            const double *row = partial + (size_t)t * n_cols;
            for (int32_t c = c0; c < c1; ++c) acc += row[map_col[c]];
        }
    }
    *out = acc;
}

}  // namespace

template <int kSpl>
const void *reverse_symbol(bool global_slots) {
    return global_slots ? reinterpret_cast<const void *>(&reverse_kernel<true, kSpl>)
                        : reinterpret_cast<const void *>(&reverse_kernel<false, kSpl>);
}

const void *reverse_symbol(bool global_slots, int spl) {
    return spl == 2 ? reverse_symbol<2>(global_slots) : reverse_symbol<1>(global_slots);
}

hipError_t launch_reverse(const RevParams &p, int spl, int grid, size_t lds_bytes, hipStream_t stream) {
    const bool gl = p.global_slots > 0;
    const dim3 block((unsigned)(p.waves * kLanes));
    if (spl == 2) {
        if (gl)
            hipLaunchKernelGGL((reverse_kernel<true, 2>), dim3((unsigned)grid), block, lds_bytes, stream, p);
        else
            hipLaunchKernelGGL((reverse_kernel<false, 2>), dim3((unsigned)grid), block, lds_bytes, stream, p);
    } else {
        if (gl)
            hipLaunchKernelGGL((reverse_kernel<true, 1>), dim3((unsigned)grid), block, lds_bytes, stream, p);
        else
            hipLaunchKernelGGL((reverse_kernel<false, 1>), dim3((unsigned)grid), block, lds_bytes, stream, p);
    }
    return hipGetLastError();
}

hipError_t launch_unit_draws(const Record *links, int32_t n_links, const int32_t *seeds, int32_t first, int32_t count,
                             int32_t n, const DistParams *dists, const AliasRec *alias, double *out, int32_t ld_out,
                             hipStream_t stream) {
    if (n_links <= 0 || count <= 0) return hipSuccess;
    constexpr int32_t kMaxRows = 65535;  // grid y dimension per launch
    for (int32_t q0 = 0; q0 < n_links; q0 += kMaxRows) {
        const int32_t m = std::min(kMaxRows, n_links - q0);
        hipLaunchKernelGGL(unit_draw_kernel, dim3((unsigned)((count + 255) / 256), (unsigned)m), dim3(256), 0, stream,
                           links + q0, m, seeds, first, count, n, dists, alias, out + (size_t)q0 * (uint32_t)ld_out,
                           (uint32_t)ld_out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t reverse_occupancy(int waves, size_t lds_bytes, bool global_slots, int spl, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, reverse_symbol(global_slots, spl),
                                                        waves * kLanes, lds_bytes);
}

hipError_t launch_reduce_columns(const double *partial, int32_t n_tiles, int32_t n_cols, double *groups,
                                 const int32_t *map_off, const int32_t *map_col, int32_t L, int32_t E, int accumulate,
                                 double *grad_dmin, double *grad_earliest, double *objective, hipStream_t stream) {
    const int32_t n_groups = (n_tiles + kReduceGroup - 1) / kReduceGroup;
    if (n_groups > 0) {
        hipLaunchKernelGGL(group_sum_kernel, dim3((unsigned)((n_cols + 255) / 256), (unsigned)n_groups), dim3(256), 0,
                           stream, partial, n_tiles, n_cols, groups);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const int32_t n_out = L + E + 1;
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, stream, groups,
                       n_groups, n_cols, map_off, map_col, L, E, accumulate, grad_dmin, grad_earliest, objective);
    return hipGetLastError();
}

}  // namespace mcdp
