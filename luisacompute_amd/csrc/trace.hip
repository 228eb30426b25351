// trace.hip -- software trace_closest / trace_any on gfx950 (no ray-tracing hardware on CDNA4).
//
// Replaces the reference's lc_accel_trace_closest / lc_accel_trace_any
// (src/backends/cuda/cuda_builtin/cuda_device_resource.h:1783-1829, OptiX hardware traversal)
// with the observable semantics of the CPU backend's AccelImpl::trace_closest / trace_any
// (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:449-535); the per-ray traversal is the
// Walker of traverse.h.
//
// Launch design (wave64): a persistent grid; a wave hands rays to its idle lanes from a pool of
// queue slots it claims in chunks (per-XCD queue segments, below), runs `steps` traversal steps,
// then refills finished lanes, so lanes stay busy while rays of different lengths finish. The queue is the concatenation of up to kMaxBatches ray
// batches (each with its own output buffer and visibility mask); an optional permutation
// (`order`) makes lanes fetch rays in sorted order. The ray count is a kernel argument, or is
// read from device memory (`count_ptr`) for launches whose queue length is produced on the GPU.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <map>
#include <mutex>

#include "lc_hip_bvh.h"
#include "lc_hip_internal.h"
#include "traverse.h"

// Working memory of lc_hip_accel_trace_closest_noalloc: the launch's slot counters, error word
// and stats in the first kNoAllocWordsBytes, then the per-lane stack spill area. Each GPU holds
// its own copy for the life of the process; host code reaches it only through the address that
// noalloc_scratch_lookup returns for the current device.
static __device__ __align__(256) unsigned char scratch[lchip::kNoAllocScratchBytes];

namespace lchip {

// Address of the current device's copy of `scratch`, looked up on the first call for each device
// (with that device current) and kept per device index for later calls.
hipError_t noalloc_scratch_lookup(void **device_ptr) {
    static std::mutex mu;
    static std::map<int, void *> by_device;
    *device_ptr = nullptr;
    int index = 0;
    hipError_t e = hipGetDevice(&index);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    auto it = by_device.find(index);
    if (it != by_device.end()) {
        *device_ptr = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    by_device.emplace(index, p);
    *device_ptr = p;
    return hipSuccess;
}

// Template parameters: ANY (trace_any), TWO (TLAS + BLAS; otherwise one instance's BLAS),
// COUNT (count node visits / triangle tests), BLOCK threads, SLDS stack entries in LDS.
// With the 8-entry LDS stack the single-level forms fit 7 (any-hit, <= 72 VGPRs) and 6
// (closest-hit, <= 80 VGPRs) waves per SIMD without spilling; the hint keeps the register
// allocator there (deeper LDS stacks are limited by LDS first).
template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((!TWO && SLDS == 8) ? (ANY ? 7 : 6) : 1)))
void k_trace(const TraceParams P) {
    __shared__ uint2 stk[SLDS > 0 ? SLDS : 1][BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint64_t glane = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint32_t count = P.count_ptr ? min(*P.count_ptr, P.count) : P.count;
    trav::Walker<ANY, TWO, COUNT, BLOCK, SLDS> W;
    W.init(P.geo, stk, tid, P.spill + glane * (uint64_t)P.spill_depth, P.spill_depth, P.error);

    // Fetching: the queue is split into kFetchSegments contiguous segments, one per XCD
    // (workgroups are dealt to the XCDs round-robin, so block b runs on XCD b mod 8), each with a
    // slot counter on its own cache line. A wave claims `chunk` slots of its segment per atomic
    // and hands them to its idle lanes from that pool; when the segment is used up it takes
    // slots from the others. One counter address serialises its atomics (about 85M/s), so a
    // per-round atomic on one counter would cap the launch; a segment per XCD also keeps each
    // XCD's L2 on one region of a sorted queue.
    // fetch_chunk 0 claims exactly the slots the idle lanes need (the narrowest window of the
    // queue in flight, one atomic per refill round); fetch_segments 1 or kFetchSegments.
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t chunk = P.fetch_chunk ? max(64u, min(P.fetch_chunk, (count / max(1u, waves * 4u)) & ~63u)) : 0u;
    const uint32_t nseg = P.fetch_segments > 1 ? kFetchSegments : 1u;
    const uint32_t seg0 = blockIdx.x % nseg;
    const bool simple = chunk == 0u && nseg == 1u;
    uint32_t pool = 0, pool_end = 0; // wave-uniform: unclaimed slots [pool, pool_end)
    bool drained = false;            // wave-uniform: every segment is used up
    bool active = false, exhausted = false;
    uint64_t rounds = 0; // counting form: refill rounds of this wave
    uint32_t ray_id = 0; // index within the batch
    uint32_t batch = 0;  // batch of the current ray
    for (;;) {
        const uint64_t need = __ballot(!active && !exhausted);
        if (need) {
            const uint32_t n_need = (uint32_t)__popcll(need);
            const uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            const uint32_t avail = pool_end - pool;
            uint32_t my = rank < avail ? pool + rank : 0xFFFFFFFFu; // lanes served by the pool
            if (simple) {
                // one counter, exactly the slots the idle lanes need
                const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                uint32_t b = 0;
                if (lane == leader) b = atomicAdd(P.ray_counter, n_need);
                b = __builtin_amdgcn_readlane(b, (int)leader);
                my = b + rank;
                drained = b + n_need >= count;
            } else if (avail >= n_need) {
                pool += n_need;
            } else if (drained) {
                pool = pool_end;
            } else {
                // refill: the leader claims a chunk of this XCD's segment, else of another one
                const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                uint32_t nb = 0, ne = 0;
                if (lane == leader) {
                    const uint32_t claim = chunk ? chunk : n_need - avail;
                    for (uint32_t k = 0; k < nseg; k++) {
                        const uint32_t s = (seg0 + k) % nseg;
                        const uint32_t lo = (uint32_t)((uint64_t)count * s / nseg);
                        const uint32_t hi = (uint32_t)((uint64_t)count * (s + 1) / nseg);
                        if (lo == hi) continue;
                        const uint32_t b = atomicAdd(P.ray_counter + s * kCounterStride, claim);
                        if (b < hi - lo) {
                            nb = lo + b;
                            ne = lo + min(b + claim, hi - lo);
                            break;
                        }
                    }
                }
                nb = __builtin_amdgcn_readlane(nb, (int)leader); // wave-uniform from here on
                ne = __builtin_amdgcn_readlane(ne, (int)leader);
                drained = nb == ne;
                if (rank >= avail && rank - avail < ne - nb) my = nb + (rank - avail);
                pool = nb + min(n_need - avail, ne - nb);
                pool_end = ne;
            }
            if (!active && !exhausted) {
                if (my < count) {
                    const uint32_t q = P.order ? P.order[my] : my; // sorted queue: fetch in key order
                    batch = find_batch(P.batches, q);
                    ray_id = q - P.batches.b[batch].start;
                    const float4 *rays = P.batches.b[batch].rays;
                    const float4 r0 = rays[2 * (uint64_t)ray_id], r1 = rays[2 * (uint64_t)ray_id + 1];
                    W.start(P.geo, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, P.batches.b[batch].mask);
                    active = true;
                } else if (drained) {
                    exhausted = true; // no slot left in any segment (else: retry next round)
                }
            }
        }
        if (!__any(active)) break;
        if (COUNT && lane == (uint32_t)__ffsll((long long)__ballot(1)) - 1u) rounds++;
        if (active && W.run(P.geo, P.steps, P.tri_quarters, P.refill_below, !drained)) {
            // the counting form has no output buffers (out is null)
            if (COUNT) {
            } else if (ANY) {
                static_cast<uint8_t *>(P.batches.b[batch].out)[ray_id] = W.any_hit ? 1u : 0u;
            } else {
                W.write_hit(reinterpret_cast<uint2 *>(static_cast<char *>(P.batches.b[batch].out) + 24ull * ray_id),
                            P.hit_record != 0u);
            }
            active = false;
        }
    }
    if (COUNT) {
        // stats: node visits, triangle tests, live lane-iterations, wave iterations, refill rounds
        uint64_t v[5] = {W.n_nodes, W.n_tris, W.n_live, W.n_wave, rounds};
        for (int k = 0; k < 5; k++) {
            for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
            if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long *>(P.stats) + k, (unsigned long long)v[k]);
        }
    }
}

__global__ void k_zero_counter(uint32_t *c) {
    if (threadIdx.x < kFetchSegments) c[threadIdx.x * kCounterStride] = 0u;
}

namespace {

template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
hipError_t launch_one(const TraceParams &P, hipStream_t stream, int grid) {
    hipLaunchKernelGGL((k_trace<ANY, TWO, COUNT, BLOCK, SLDS>), dim3(grid), dim3(BLOCK), 0, stream, P);
    return hipGetLastError();
}

template<bool ANY, bool TWO, bool COUNT>
hipError_t launch_cfg(const TraceParams &P, hipStream_t stream, int grid, int block, int slds) {
#define LC_TRACE_CFG(B, S) \
    if (block == B && slds == S) return launch_one<ANY, TWO, COUNT, B, S>(P, stream, grid);
    LC_TRACE_CFG(64, 8) LC_TRACE_CFG(64, 16) LC_TRACE_CFG(64, 24)
    LC_TRACE_CFG(128, 8) LC_TRACE_CFG(128, 16) LC_TRACE_CFG(128, 24)
    LC_TRACE_CFG(256, 8) LC_TRACE_CFG(256, 16) LC_TRACE_CFG(256, 24)
#undef LC_TRACE_CFG
    return hipErrorInvalidValue;
}

template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
int occupancy_one() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_trace<ANY, TWO, COUNT, BLOCK, SLDS>, BLOCK, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template<bool ANY, bool TWO, bool COUNT>
int occupancy_cfg(int block, int slds) {
#define LC_OCC_CFG(B, S) \
    if (block == B && slds == S) return occupancy_one<ANY, TWO, COUNT, B, S>();
    LC_OCC_CFG(64, 8) LC_OCC_CFG(64, 16) LC_OCC_CFG(64, 24)
    LC_OCC_CFG(128, 8) LC_OCC_CFG(128, 16) LC_OCC_CFG(128, 24)
    LC_OCC_CFG(256, 8) LC_OCC_CFG(256, 16) LC_OCC_CFG(256, 24)
#undef LC_OCC_CFG
    return 1;
}

} // namespace

bool trace_config_supported(int block, int slds) {
    return (block == 64 || block == 128 || block == 256) && (slds == 8 || slds == 16 || slds == 24);
}

int trace_blocks_per_cu(bool any, bool two, bool count, int block, int slds) {
    if (any) {
        if (two) return count ? occupancy_cfg<true, true, true>(block, slds) : occupancy_cfg<true, true, false>(block, slds);
        return count ? occupancy_cfg<true, false, true>(block, slds) : occupancy_cfg<true, false, false>(block, slds);
    }
    if (two) return count ? occupancy_cfg<false, true, true>(block, slds) : occupancy_cfg<false, true, false>(block, slds);
    return count ? occupancy_cfg<false, false, true>(block, slds) : occupancy_cfg<false, false, false>(block, slds);
}

hipError_t trace_launch(const TraceParams &P, hipStream_t stream, bool any, bool two, bool count, int grid, int block,
                        int slds) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, stream, P.ray_counter);
    if (any) {
        if (two) return count ? launch_cfg<true, true, true>(P, stream, grid, block, slds) : launch_cfg<true, true, false>(P, stream, grid, block, slds);
        return count ? launch_cfg<true, false, true>(P, stream, grid, block, slds) : launch_cfg<true, false, false>(P, stream, grid, block, slds);
    }
    if (two) return count ? launch_cfg<false, true, true>(P, stream, grid, block, slds) : launch_cfg<false, true, false>(P, stream, grid, block, slds);
    return count ? launch_cfg<false, false, true>(P, stream, grid, block, slds) : launch_cfg<false, false, false>(P, stream, grid, block, slds);
}

} // namespace lchip
