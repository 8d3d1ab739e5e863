// gfx950 kernels of the Monte Carlo DAG propagation path.
//
// walk_kernel fuses the two hot loops of Simulator::run (_core.cpp:323-329
// delay sampling, :332-350 max-plus propagation) for 64 samples per
// workgroup, one sample per lane:
//   * plan records (events, predecessors, distributions) are wave-uniform and
//     arrive through scalar loads; the distribution switch is a scalar branch;
//   * realized times of predecessors that later steps read live in LDS
//     (slot-major, 64 doubles per slot, so every access is one conflict-free
//     ds_read_b64 / ds_write_b64 per lane);
//   * outputs are sample-contiguous rows, so each wave stores 512-byte runs.
// The arithmetic of propagation is FP64 add and compare-select only, in the
// reference's order: t = realized[src] + dur; t = (upper < t) ? upper : t;
// if (t >= best) {best = t; cause = src}; realized = (upper < best) ? upper : best.
// std::min is (b < a) ? b : a, so a NaN candidate is never taken; no v_min_f64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcdp_kernels.hpp"
#include "mcdp_sampler.hpp"

namespace mcdp {

template <bool kInjected>
__global__ void __launch_bounds__(kMaxWaves * kLanes) walk_kernel(WalkParams p) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = p.waves;
    const int64_t ld = p.ld;
    // The plan never changes during a launch: reading it through the constant
    // address space lets wave-uniform record loads become scalar loads.
    const auto step_off = as_const_space(p.step_off);
    const auto evs = as_const_space(p.ev);
    const auto prs = as_const_space(p.pr);
    const auto dists = as_const_space(p.dist);
    const auto alias = as_const_space(p.alias);

    for (int64_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x) {
        const int64_t s = tile * kLanes + lane;
        const bool active = s < p.n;
        SeedKey key{0u, 0u};
        if (!kInjected) key = seed_key(active ? p.seeds[s] : 0);

        for (int t = 0; t < p.n_steps; ++t) {
            const int b = step_off[t * W + wave];
            const int e = step_off[t * W + wave + 1];
            double res[kMaxEventsPerWave];
            for (int i = b; i < e; ++i) {
                const EventRec er = evs[i];
                double best = er.earliest;
                int32_t who = -1;
                for (int k = er.pred_begin; k < er.pred_begin + er.pred_count; ++k) {
                    const PredRec pr = prs[k];
                    double rs;
                    if (pr.src_slot >= 0) {
                        rs = lds[pr.src_slot * kLanes + lane];
                    } else if (pr.src_slot == kSrcGlobal) {
                        rs = active ? p.realized[(int64_t)pr.src_event * ld + s] : 0.0;
                    } else {
                        rs = pr.src_value;
                    }
                    double dur = 0.0;
                    if (pr.info & kInfoEdgeValid) {
                        if (kInjected) {
                            dur = active ? __builtin_nontemporal_load(p.dur_in + (int64_t)pr.edge * ld + s) : 0.0;
                        } else {
                            const uint32_t di = pr.info & kInfoDistMask;
                            dur = (di == kNoDist) ? pr.dmin
                                                  : draw_duration(dists[di], alias, (uint32_t)pr.edge, pr.dmin, key);
                            if ((pr.info & kInfoOwner) && p.durations && active)
                                __builtin_nontemporal_store(dur, p.durations + (int64_t)pr.edge * ld + s);
                        }
                    }
                    double tt = rs + dur;
                    tt = (er.upper < tt) ? er.upper : tt;
                    if (tt >= best) {
                        best = tt;
                        who = pr.src_event;
                    }
                }
                const double r = (er.upper < best) ? er.upper : best;
                if (active) {
                    if (p.realized) __builtin_nontemporal_store(r, p.realized + (int64_t)er.event * ld + s);
                    if (p.cause) __builtin_nontemporal_store(who, p.cause + (int64_t)er.event * ld + s);
                }
                res[i - b] = r;
            }
            // Every read of this step is done before any slot is overwritten,
            // and every commit is visible before the next step reads it.
            __syncthreads();
            for (int i = b; i < e; ++i) {
                const int slot = evs[i].slot;
                if (slot >= 0) lds[slot * kLanes + lane] = res[i - b];
            }
            __syncthreads();
        }

        // Activities no predecessor entry references still get a duration row.
        if (!kInjected && p.durations) {
            const auto orph = as_const_space(p.orph);
            for (int j = wave; j < p.n_orphans; j += W) {
                const OrphanRec o = orph[j];
                const uint32_t di = o.info & kInfoDistMask;
                const double dur =
                    (di == kNoDist) ? o.dmin : draw_duration(dists[di], alias, (uint32_t)o.edge, o.dmin, key);
                if (active) __builtin_nontemporal_store(dur, p.durations + (int64_t)o.edge * ld + s);
            }
        }
    }
}

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

hipError_t launch_walk(const WalkParams &p, bool injected, int grid, size_t lds_bytes, hipStream_t stream) {
    const dim3 block(p.waves * kLanes);
    if (injected)
        hipLaunchKernelGGL(walk_kernel<true>, dim3(grid), block, lds_bytes, stream, p);
    else
        hipLaunchKernelGGL(walk_kernel<false>, dim3(grid), block, lds_bytes, stream, p);
    return hipGetLastError();
}

hipError_t walk_occupancy(bool injected, int waves, size_t lds_bytes, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, injected ? (const void *)walk_kernel<true> : (const void *)walk_kernel<false>,
        waves * kLanes, lds_bytes);
}

hipError_t launch_philox_debug(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out,
                               hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    if (grid > 0) hipLaunchKernelGGL(philox_debug_kernel, dim3(grid), dim3(block), 0, stream, variant, n, ctr, key, out);
    return hipGetLastError();
}

}  // namespace mcdp
