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
//                i2 = number of successor entries,
//                info = kInfoHeader | number of predecessor records that follow;
//   successors:  not records but (slot, position) pairs in `succ`, from
//                ev_succ: per target event t that lists this event as a
//                predecessor, t's slot and this event's position in t's list
//                (one pair per position: a source listed twice has two);
//   predecessor: in the event's input order; i0 = source, i1 = column of the
//                minimal-duration derivative (-1: none; for a stale entry the
//                first of two columns, the second for the source's earliest
//                time; columns are numbered in the order a wave writes them,
//                from col_off of its step), i2 = edge (durations row), a = 1 / d (kMulRead) or 1
//                (kMulDraw), info = the forward record's kind, flags and
//                distribution index (kMulDraw: unit row) with the bits below.
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
    kMulDraw = 3,  // duration = d + X d, d == 0: 1 + X, drawn again from the seed into a row of
                   // unit durations (record bits 13..30: the row, a = 1)
};

constexpr int kRevEventsPerWave = 8;  // events of one wave and step (compile-time bound)

// The reverse walk's per-event blocks, one stream per (step, wave), in dwords:
//   [0] event  [1] slot (-1: none)  [2] predecessors | successors << 16  [3] dwords of the block
//   [4..5] earliest  [6..7] upper (earliest + max_delay, rounded once)
//   per predecessor, in the event's input order (kBlkPred dwords): [0..1] a (1 / d for kMulRead,
//     1 otherwise), [2] source, [3] edge, [4] info (the forward record's kind and flags, kRevClass,
//     kRevDup, kRevStale; kMulDraw: the unit row in the distribution-index bits)
//   per successor (kBlkSucc dwords): the target's slot, this event's position in the target's list
// A wave loads a block's first kBlkWindow dwords with one vector load (a dword per lane) and reads
// fields with readlane; the stream ends in kBlkWindow padding dwords.
constexpr int kBlkHeader = 8, kBlkPred = 5, kBlkSucc = 2, kBlkWindow = 64;

struct RevParams {
    const uint32_t *blocks;   // the block streams
    const int32_t *blk_off;   // [n_steps * waves + 1]: first dword of each (step, wave) stream
    const int32_t *ev_off;    // [n_steps * waves + 1]: event range of (step, wave)
    const int32_t *col_off;   // [n_steps * waves]: first partial-row column (step, wave) writes
    const int32_t *ev_slot;   // slots of the events in stream order (-1: none)
    const DistParams *dists;  // the forward plan's distributions and alias tables (kMulDraw)
    const AliasRec *alias;
    const double *earliest;   // [E] (repeated sources read stale predecessors' earliest time)
    int32_t n_steps;
    int32_t waves;
    int32_t lds_slots;        // slots in LDS: 64 spl adjoints (double) + 64 spl winner positions (int32) each
    int32_t global_slots;     // further slots in global memory, per workgroup
    int32_t L;                // links
    int32_t n_cols;           // columns of a partial row
    int32_t obj_col;          // first of `waves` objective columns
    int32_t tile0;            // first tile (of 64 spl samples) of this launch
    const double *weights;    // [E]
    const double *realized;   // [E][ld]
    const double *durations;  // [L][ld]
    const int32_t *cause;     // [E][ld]
    const double *unit;       // [unit links][ld_unit]: unit durations for this launch's tiles (kMulDraw)
    int32_t ld_unit;
    int32_t n;                // samples of the whole call
    int32_t ld;
    int32_t n_tiles;          // tiles of this launch
    double *partial;          // [n_tiles][n_cols]
    double *gslots;           // per workgroup: [global_slots][64 spl] adjoints, then [global_slots][64 spl]
                              // int32 positions, in 2 * global_slots * 64 spl doubles
};

// spl: samples per lane (1 or 2), so a tile is 64 spl samples.
hipError_t launch_reverse(const RevParams &p, int spl, int grid, size_t lds_bytes, hipStream_t stream);
// Unit durations of `n_links` links (their forward records) for samples
// first .. first + count - 1 (clamped to n - 1): out[q * ld_out + i].
hipError_t launch_unit_draws(const Record *links, int32_t n_links, const int32_t *seeds, int32_t first, int32_t count,
                             int32_t n, const DistParams *dists, const AliasRec *alias, double *out, int32_t ld_out,
                             hipStream_t stream);
hipError_t reverse_occupancy(int waves, size_t lds_bytes, bool global_slots, int spl, int *blocks_per_cu);
// Tiles per group of the two-stage column reduction (grid dimension y of the
// first stage: at most 65535 groups per launch).
constexpr int kReduceGroup = 16;
// groups[g][c] = sum of partial[t][c] over the kReduceGroup tiles of group g,
// in order; then out_X[o] (+)= sum over groups g, in order, of sum over
// map_col[map_off[o] .. map_off[o + 1]) of groups[g][col].  Outputs o < L go to
// grad_dmin, the next E to grad_earliest, the last one to objective.  `groups`
// holds ceil(n_tiles / kReduceGroup) rows of n_cols.
hipError_t launch_reduce_columns(const double *partial, int32_t n_tiles, int32_t n_cols, double *groups,
                                 const int32_t *map_off, const int32_t *map_col, int32_t L, int32_t E, int accumulate,
                                 double *grad_dmin, double *grad_earliest, double *objective, hipStream_t stream);

}  // namespace mcdp
