// Launch interface between the host compiler (mcdp_host.cpp) and the gfx950
// kernels (mcdp_kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "mcdp_plan.hpp"

namespace mcdp {

constexpr int kMaxWaves = 16;  // waves per workgroup (1024 threads)

// Events one wave evaluates in one step: a compile-time bound, because each
// wave keeps its step's results in registers until the commit (beyond 16 the
// compiler moves that array to scratch memory).  The host builds the schedule
// for one of these bounds.
constexpr int kEventsPerWaveChoices[] = {4, 8, 12, 16};
constexpr int kMaxEventsPerWave = 16;

// Samples per lane: a workgroup's tile is 64 x spl samples.  Two samples per
// lane halve the scalar work (record loads, distribution dispatch) per sample
// and give each wave two independent draw chains; they double the LDS of a
// slot and the registers of the step results, so they pair with 4 or 8
// events per wave.
constexpr int kSplChoices[] = {1, 2};
constexpr int kMaxEventsPerWaveSpl2 = 8;

// Alias tables up to this size are staged in LDS once per workgroup.
constexpr size_t kAliasLdsMaxBytes = 16 * 1024;

// Kernel variants: the fast one assumes every predecessor the plan reads is
// in LDS (or stale) and the alias tables fit in LDS, so the predecessor loop
// issues no vector memory load (a vector load there would also wait for all
// of the wave's outstanding output stores: gfx9 counts both in vmcnt).
enum WalkPaths : int { kPathsFast = 0, kPathsGeneral = 1 };

struct WalkParams {
    const int32_t *step_off;   // [n_steps * waves + 1]: record range of (step, wave)
    const Record *recs;        // record streams, ending with one padding record
    const int32_t *ev_off;     // [n_steps * waves + 1]: event range of (step, wave) in slot_list
    const int32_t *slot_list;  // LDS slot of each event in stream order (-1: none)
    const int32_t *step_slots; // [n_steps * waves][k]: the slots of (step, wave)'s events, -1 past them
    const Record *orph;        // activities no predecessor entry references
    const DistParams *dists;   // compiled distributions (p1..p4 of the rarer kinds)
    const AliasRec *alias;
    const double *earliest;    // [E]: realized value of a stale predecessor
    const int32_t *ev_rec;     // header record of each event, in slot_list order (wide steps)
    int32_t n_steps;
    int32_t n_orphans;
    int32_t waves;
    int32_t n_alias;          // AliasRecs (staged to LDS after the slots in the fast variant)
    int32_t slots;            // LDS predecessor slots
    const int32_t *seeds;   // sampled mode: int32[n]
    const double *dur_in;   // injected mode: [L][ld]
    int32_t n;              // samples (< 2^31)
    int32_t ld;             // leading dimension of every per-sample array (< 2^31)
    int32_t n_tiles;        // ceil(n / (64 spl))
    double *realized;       // [E][ld] (read within the launch when a plan spills slots)
    double *durations;      // [L][ld] or null (sampled mode)
    int32_t *cause;         // [E][ld] or null
    uint8_t *cause_pos;     // [E][ld] or null: the winning entry's position in the event's list (0xFF: none),
                            // instead of cause, for events of at most 254 entries
};

// A batch: several plans' WalkParams and a queue of (plan, tile) items.
struct WalkBatchParams {
    const WalkParams *params;  // [plans]
    const int32_t *item_dag;   // [n_items] plan of each item
    const int32_t *item_tile;  // [n_items] tile of each item
    int32_t n_items;
    int32_t *counter;          // queue head, zeroed before the launch
};

struct WalkVariant {
    bool injected;  // durations read from dur_in instead of sampled
    int paths;      // WalkPaths
    int k;          // one of kEventsPerWaveChoices
    int spl;        // samples per lane: one of kSplChoices (spl 2: k <= kMaxEventsPerWaveSpl2)
    bool pos = false;  // sampled walks only: the cause output is WalkParams::cause_pos
};

hipError_t launch_walk(const WalkParams &p, WalkVariant v, int grid, size_t lds_bytes, hipStream_t stream);
hipError_t walk_occupancy(WalkVariant v, int waves, size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_walk_batch(const WalkBatchParams &b, WalkVariant v, int waves, int grid, size_t lds_bytes,
                             hipStream_t stream);
hipError_t walk_batch_occupancy(WalkVariant v, int waves, size_t lds_bytes, int *blocks_per_cu);
// Predecessor records a wide-step wave loads together; the record array ends
// in this many padding records.
constexpr int kWideBatch = 4;

// One step of the walk spread over the device (few samples): events e0 ..
// e0 + n_events - 1 of the schedule, `groups` groups of 64 samples; then the
// orphan activities.  Steps are separated by stream order.
hipError_t launch_walk_wide(const WalkParams &p, bool injected, int32_t e0, int32_t n_events, int32_t groups,
                            hipStream_t stream);
hipError_t launch_walk_wide_orphans(const WalkParams &p, int32_t groups, hipStream_t stream);
hipError_t launch_copy(double *dst, const double *src, int64_t n, hipStream_t stream);
hipError_t launch_math_debug(int fn, int64_t n, const double *x, double *out, hipStream_t stream);
hipError_t launch_philox_debug(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out,
                               hipStream_t stream);

}  // namespace mcdp
