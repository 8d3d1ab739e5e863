// Launch interface of the reverse (pathwise-gradient) walk, mcdp_grad.hip.
//
// The forward walk evaluates realized = max-plus over predecessors; its
// derivative, where it exists, follows one entry per event: the one that won
// the max (or none, when the event kept its earliest time or hit its cap).
// The reverse walk visits the forward plan's levels in reverse order with the
// samples on lanes, as the forward walk does, and carries per sample the
// adjoint g_e = dJ/d realized_e.  An event gathers its adjoint from its
// successors' LDS slots (each successor stores its adjoint and the position
// of its winning entry), then passes it on through its own slot.  Per-link and
// per-event derivatives are summed over a tile's 64 samples with a wave
// reduction and written to that tile's row of a partial-sum matrix, which a
// second kernel reduces over tiles in a fixed order: results are
// deterministic and do not depend on how samples are split into launches.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "mcdp_plan.hpp"

namespace mcdp {

// Reverse-walk records (the 32-byte Record of the forward plan).
//   header:      a = earliest, b = earliest + max_delay, i0 = event,
//                i1 = its slot (-1: no predecessor reads its adjoint),
//                i2 = number of successor records that follow,
//                info = kInfoHeader | number of predecessor records after those;
//   successor:   i0 = a target event t that lists this event as a predecessor,
//                i1 = t's slot, i2 = the position of this event in t's list
//                (one record per position: a source listed twice has two);
//   predecessor: in the event's input order; i0 = source, i1 = column of the
//                minimal-duration derivative (-1: none; for a stale entry the
//                first of two columns, the second for the source's earliest
//                time), i2 = edge (durations row and draw key), a = 1 / d
//                (kMulRead), b = p0 of the distribution (kMulDraw), info =
//                the forward record's kind, flags and distribution index with
//                the bits below.
enum : uint32_t {
    kRevClassShift = 4,
    kRevClassMask = 3u << kRevClassShift,  // d duration / d minimal duration:
    kRevDup = 1u << 6,                     // the source occurs more than once in the list
    kRevStale = 1u << 11,                  // read before the source is evaluated: earliest[src]
    kRevNpredMask = 0x7FFFFFFFu,           // header: predecessor records
};
enum RevMul : uint32_t {
    kMulNone = 0,  // edge outside the link table: no minimal duration
    kMulOne = 1,   // duration = d or d + X
    kMulRead = 2,  // duration = d + X d, d != 0: duration * (1 / d)
    kMulDraw = 3,  // duration = d + X d, d == 0: 1 + X, drawn again from the seed
};

constexpr int kRevEventsPerWave = 8;  // events of one wave and step (compile-time bound)

struct RevParams {
    const int32_t *step_off;  // [n_steps * waves + 1]: record range of (step, wave)
    const Record *recs;
    const int32_t *ev_off;    // [n_steps * waves + 1]: range of (step, wave) in ev_list
    const int32_t *ev_list;   // events in stream order
    const int32_t *ev_slot;   // their slots (-1: none)
    const DistParams *dists;  // the forward plan's distributions and alias tables (kMulDraw)
    const AliasRec *alias;
    const double *earliest;   // [E] (repeated sources read stale predecessors' earliest time)
    int32_t n_steps;
    int32_t waves;
    int32_t lds_slots;        // slots in LDS: 64 adjoints (double) + 64 winner positions (int32) each
    int32_t global_slots;     // further slots in global memory, per workgroup
    int32_t L;                // column of event e's earliest-time derivative: L + e
    int32_t n_cols;           // columns of a partial row
    int32_t obj_col;          // first of `waves` objective columns
    int32_t tile0;            // first tile of this launch
    const int32_t *seeds;     // [n] (kMulDraw entries only)
    const double *weights;    // [E]
    const double *realized;   // [E][ld]
    const double *durations;  // [L][ld]
    const int32_t *cause;     // [E][ld]
    int32_t n;                // samples of the whole call
    int32_t ld;
    int32_t n_tiles;          // tiles of this launch
    double *partial;          // [n_tiles][n_cols]
    double *gslots;           // per workgroup: [global_slots][64] adjoints, then [global_slots][64] int32
                              // positions, in 2 * global_slots * 64 doubles
};

hipError_t launch_reverse(const RevParams &p, int grid, size_t lds_bytes, hipStream_t stream);
hipError_t reverse_occupancy(int waves, size_t lds_bytes, int *blocks_per_cu);
// out_X[o] (+)= sum over tiles t, in order, of sum over map_col[map_off[o] ..
// map_off[o + 1]) of partial[t][col]; outputs o < L go to grad_dmin, the next
// E to grad_earliest, the last one to objective.
hipError_t launch_reduce_columns(const double *partial, int32_t n_tiles, int32_t n_cols, const int32_t *map_off,
                                 const int32_t *map_col, int32_t L, int32_t E, int accumulate, double *grad_dmin,
                                 double *grad_earliest, double *objective, hipStream_t stream);

}  // namespace mcdp
