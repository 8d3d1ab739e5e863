// Launch interface between the host compiler (mcdp_host.cpp) and the gfx950
// kernels (mcdp_kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "mcdp_plan.hpp"

namespace mcdp {

constexpr int kMaxWaves = 16;           // waves per workgroup (1024 threads)
constexpr int kMaxEventsPerWave = 16;   // events one wave evaluates in one step

struct WalkParams {
    const int32_t *step_off;  // [n_steps * waves + 1]
    const EventRec *ev;
    const PredRec *pr;
    const OrphanRec *orph;
    const DistRec *dist;
    const AliasRec *alias;
    int32_t n_steps;
    int32_t n_orphans;
    int32_t waves;
    int32_t pad;
    const int32_t *seeds;   // sampled mode: int32[n]
    const double *dur_in;   // injected mode: [L][ld]
    int64_t n;              // samples
    int64_t ld;             // leading dimension of every per-sample array
    int64_t n_tiles;        // ceil(n / 64)
    double *realized;       // [E][ld] or null
    double *durations;      // [L][ld] or null (sampled mode)
    int32_t *cause;         // [E][ld] or null
};

hipError_t launch_walk(const WalkParams &p, bool injected, int grid, size_t lds_bytes, hipStream_t stream);
hipError_t walk_occupancy(bool injected, int waves, size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_copy(double *dst, const double *src, int64_t n, hipStream_t stream);
hipError_t launch_philox_debug(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out,
                               hipStream_t stream);

}  // namespace mcdp
