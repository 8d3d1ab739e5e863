// gfx950 kernels of the Monte Carlo DAG propagation path.
//
// walk_kernel fuses the two hot loops of Simulator::run (_core.cpp:323-329
// delay sampling, :332-350 max-plus propagation) for 64 samples per
// workgroup, one sample per lane:
//   * plan records (events, predecessors with their distribution inline) are
//     wave-uniform and arrive through scalar loads, the next one prefetched
//     while the current one is sampled; the distribution switch is a scalar
//     branch;
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

#include <algorithm>

#include "mcdp_kernels.hpp"
#include "mcdp_sampler.hpp"

namespace mcdp {

template <bool kInjected, int kPaths, int K>
__global__ void __launch_bounds__(kMaxWaves * kLanes) walk_kernel(WalkParams p) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = p.waves;
    const int64_t ld = p.ld;
    // The plan never changes during a launch: reading it through the constant
    // address space lets wave-uniform record loads become scalar loads.
    const auto step_off = as_const_space(p.step_off);
    const auto recs = as_const_space(p.recs);
    const auto ev_off = as_const_space(p.ev_off);
    const auto slot_list = as_const_space(p.slot_list);
    const auto dists = as_const_space(p.dists);
    const auto earliest = as_const_space(p.earliest);

    // Fast variant: stage the alias tables in LDS after the slots, so an
    // empirical draw is a ds_read rather than a vector memory load.
    AliasRec *const lds_alias = reinterpret_cast<AliasRec *>(lds + (size_t)p.slots * kLanes);
    if (!kInjected && kPaths == kPathsFast) {
        const auto src = as_const_space(p.alias);
        for (int i = threadIdx.x; i < p.n_alias; i += blockDim.x) lds_alias[i] = src[i];
        __syncthreads();
    }
    auto draw = [&](const Record &r, SeedKey key) -> double {
        if (kPaths == kPathsFast) return draw_duration(r.info, r.a, r.b, (uint32_t)r.i2, key, lds_alias, dists);
        return draw_duration(r.info, r.a, r.b, (uint32_t)r.i2, key, as_const_space(p.alias), dists);
    };

    for (int64_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x) {
        const int64_t s = tile * kLanes + lane;
        const bool active = s < p.n;
        SeedKey key{0u, 0u};
        if (!kInjected) key = seed_key(active ? p.seeds[s] : 0);

        for (int t = 0; t < p.n_steps; ++t) {
            const int r0 = step_off[t * W + wave];
            const int r1 = step_off[t * W + wave + 1];
            double res[K];
            int j = 0;
            if (r0 < r1) {
                // One stream of records: an event header, then its predecessors
                // in input order.  The stream ends in a padding record, so the
                // one-ahead prefetch never leaves the array.
                Record cur = recs[r0];
                double best = 0.0, upper = 0.0;
                int32_t who = -1, event = 0, remaining = -1;
                for (int r = r0; r < r1; ++r) {
                    const Record nxt = recs[r + 1];
                    if (cur.info & kInfoHeader) {
                        best = cur.a;
                        upper = cur.b;
                        event = cur.i0;
                        who = -1;
                        remaining = cur.i2;
                    } else {
                        double dur = 0.0;
                        if (cur.info & kInfoEdgeValid) {
                            if (kInjected) {
                                dur = active ? __builtin_nontemporal_load(p.dur_in + (int64_t)cur.i2 * ld + s) : 0.0;
                            } else {
                                dur = draw(cur, key);
                                if ((cur.info & kInfoOwner) && p.durations && active)
                                    __builtin_nontemporal_store(dur, p.durations + (int64_t)cur.i2 * ld + s);
                            }
                        }
                        double rs;
                        if (cur.i1 >= 0) {
                            rs = lds[cur.i1 * kLanes + lane];
                        } else if (kPaths == kPathsGeneral && cur.i1 == kSrcGlobal) {
                            rs = active ? p.realized[(int64_t)cur.i0 * ld + s] : 0.0;
                        } else {
                            rs = earliest[cur.i0];
                        }
                        double tt = rs + dur;
                        tt = (upper < tt) ? upper : tt;
                        if (tt >= best) {
                            best = tt;
                            who = cur.i0;
                        }
                        --remaining;
                    }
                    if (remaining == 0) {  // the event's last predecessor (or a header with none)
                        const double rr = (upper < best) ? upper : best;
                        if (active) {
                            __builtin_nontemporal_store(rr, p.realized + (int64_t)event * ld + s);
                            if (p.cause) __builtin_nontemporal_store(who, p.cause + (int64_t)event * ld + s);
                        }
                        res[j++] = rr;
                        remaining = -1;
                    }
                    cur = nxt;
                }
            }
            // Every read of this step is done before any slot is overwritten,
            // and every commit is visible before the next step reads it.
            __syncthreads();
            const int e0 = ev_off[t * W + wave];
            for (int i = 0; i < j; ++i) {
                const int slot = slot_list[e0 + i];
                if (slot >= 0) lds[slot * kLanes + lane] = res[i];
            }
            __syncthreads();
        }

        // Activities no predecessor entry references still get a duration row.
        if (!kInjected && p.durations) {
            const auto orph = as_const_space(p.orph);
            for (int i = wave; i < p.n_orphans; i += W) {
                const Record o = orph[i];
                const double dur = draw(o, key);
                if (active) __builtin_nontemporal_store(dur, p.durations + (int64_t)o.i2 * ld + s);
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

// The samplers' elementary functions on n inputs: 0 log_fast, 1 cos2pi_fast,
// 2 exp_fast.
__global__ void math_debug_kernel(int fn, int64_t n, const double *__restrict__ x, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    out[i] = fn == 0 ? log_fast(v) : fn == 1 ? cos2pi_fast(v) : exp_fast(v);
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

template <bool kInjected, int kPaths>
const void *walk_symbol(int k) {
    switch (k) {
        case 4: return (const void *)walk_kernel<kInjected, kPaths, 4>;
        case 8: return (const void *)walk_kernel<kInjected, kPaths, 8>;
        case 12: return (const void *)walk_kernel<kInjected, kPaths, 12>;
        case 16: return (const void *)walk_kernel<kInjected, kPaths, 16>;
        default: return nullptr;
    }
}

const void *walk_symbol(WalkVariant v) {
    if (v.injected) return v.paths == kPathsFast ? walk_symbol<true, kPathsFast>(v.k) : walk_symbol<true, kPathsGeneral>(v.k);
    return v.paths == kPathsFast ? walk_symbol<false, kPathsFast>(v.k) : walk_symbol<false, kPathsGeneral>(v.k);
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

hipError_t walk_occupancy(WalkVariant v, int waves, size_t lds_bytes, int *blocks_per_cu) {
    const void *fn = walk_symbol(v);
    if (!fn) return hipErrorInvalidValue;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, waves * kLanes, lds_bytes);
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
