// LDS-resident schedule kernels (see resident_core.hpp for the phase structure).
//
//   k_packed_init   admission of bit-packed (CB) lattices: descriptor from a
//                   staging array into the table, records zeroed, Philox +-1
//                   start (Lattice::initialize_random_spins, src/lattice.cpp:33-41)
//   k_resident      persistent grid; each workgroup pulls work items (one part
//                   of one slot: G lattices of one L) from an atomic queue and
//                   runs the part's sweeps and measurements with the lattices
//                   held in LDS; a slot's parts run in order, possibly in
//                   different workgroups (src/ising.cpp:87-149 replaced by
//                   Metropolis; src/lattice.cpp:102-120, src/mcrg.cpp:72-98 and
//                   :314-348 fused)
//   k_reset_queue   zeroes a launch's work counter and per-slot progress words
#include <mutex>

#include "resident_core.hpp"

namespace mcrg {

__global__ __launch_bounds__(256) void k_packed_init(PackedLattice *__restrict__ lats,
                                                     const PackedLattice *__restrict__ stage,
                                                     const int32_t *__restrict__ idx)
{
    const PackedLattice P = stage[blockIdx.y];
    const uint32_t L = (uint32_t)P.L;
    const uint32_t W = cb_words(L);
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0u) lats[idx[blockIdx.y]] = P;
    const uint64_t n_rec = (uint64_t)P.max_bins * record_size((uint32_t)P.n_levels);
    for (uint64_t i = j; i < n_rec; i += (uint64_t)gridDim.x * blockDim.x) P.records[i] = 0;
    if (j >= 2u * W) return;
    const uint32_t c = j / W, w = j % W;
    const uint32_t half = L >> 1;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32u; ++b) {
        const uint32_t mc = 32u * w + b;
        const uint32_t y = mc / half, k = mc % half;
        const uint32_t x = 2u * k + ((y + c) & 1u);
        const uint32_t m = y * L + x;
        const u32x4 r = philox(m >> 7, 0u, 0u, kTagInit << 28, P.k0, P.k1);
        word |= ((pick4(r, (m >> 5) & 3u) >> (m & 31u)) & 1u) << b;
    }
    P.spins[j] = word;
}

// Waits until *flag >= want (agent-scope acquire), polling with s_sleep.
// False when the 100 MHz real-time clock has advanced by more than
// kSubChunkWaitTicks first.  Not inlined: a leaf call with no stack, which
// keeps k_resident at 96 VGPRs without spilling (inlined, the loop's state
// cost three spilled registers in the whole kernel).
__device__ __noinline__ bool wait_flag(const uint32_t *flag, uint32_t want)
{
    const uint64_t t0 = (uint64_t)wall_clock64();
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
        __builtin_amdgcn_s_sleep(4);
        if ((uint64_t)wall_clock64() - t0 > kSubChunkWaitTicks) return false;
    }
    return true;
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_resident(const Slot *__restrict__ slots, uint32_t n_slots, uint32_t n_sub,
                                                 const int32_t *__restrict__ slot_lats,
                                                 PackedLattice *__restrict__ lats, uint32_t *queue, uint32_t *done,
                                                 uint32_t *err)
{
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_idx, s_ok;
    const uint32_t n_items = n_slots * n_sub;
    // The part finished last is published (done[pub_si] = pub_val) in the same
    // thread-0 block that takes the next item.  A thread-0 store of its own at
    // the end of the loop body was merged by the compiler with the next
    // iteration's thread-0 dequeue into a lane-divergent loop, in which the
    // other lanes of the first wave passed the loop-top barrier before lane 0
    // had published and dequeued, and the workgroup repeated its item forever.
    uint32_t pub_si = 0, pub_val = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            if (pub_val != 0u)
                __hip_atomic_store(done + pub_si, pub_val, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            s_idx = atomicAdd(queue, 1u);
        }
        __syncthreads();
        // the item and its slot are the same in every lane: keep them scalar
        const uint32_t idx = uni<true>(s_idx);
        __syncthreads();
        if (idx >= n_items) break;
        const uint32_t k = uni<true>(idx / n_slots), si = idx - k * n_slots;
        Slot sl = slots[si];
        const uint32_t per = uni<true>((sl.n_sweeps + n_sub - 1u) / n_sub), start = k * per;
        sl.n_sweeps = start < sl.n_sweeps ? min(per, sl.n_sweeps - start) : 0u;
        bool ok = true;
        if (k > 0u) {
            // the previous part of this slot ran in another workgroup, perhaps
            // on another XCD: wait for its write-back to be published
            if (threadIdx.x == 0) {
                const bool w = wait_flag(done + si, k);
                if (!w) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_ok = w ? 1u : 0u;
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok = uni<true>(s_ok) != 0u;
        }
        if (ok && sl.n_sweeps > 0u) {
            Team tm;
            tm.T = (uint32_t)NT / (uint32_t)sl.G;
            tm.g0 = threadIdx.x / tm.T;
            tm.gs = (uint32_t)sl.G;
            tm.u0 = threadIdx.x % tm.T;
            tm.us = tm.T;
            tm.lane = threadIdx.x;
            tm.block = tm.T > 64u;
            // Slots of either wide shape run the wide body (slot_shape), the
            // others the generic body.  Workgroups of 64 and 128 threads keep
            // the generic body only: with both, their register allocation
            // spills.
            if (NT >= 256 && slot_shape((uint32_t)sl.L, (uint32_t)sl.G, (uint32_t)sl.n_levels, (uint32_t)NT) !=
                                 kShapeGeneric)
                run_slot<kShapeRows64>(sl, slot_lats, lats, lds, tm);
            else {
                // The generic body carries the small lattices: chains of short
                // dependent steps on a few lanes per lattice, whose slots take
                // the longest in a mixed launch.  Their waves issue first on a
                // SIMD they share with wide slots, which fill the cycles left.
                __builtin_amdgcn_s_setprio(3);
                run_slot<kShapeGeneric>(sl, slot_lats, lats, lds, tm);
                __builtin_amdgcn_s_setprio(0);
            }
        }
        if (n_sub > 1u) {
            // this part's write-back (spins, records, counters) goes to the
            // workgroup that runs the next part: released here by every
            // thread, published at the loop top; after a timed-out wait the
            // part is skipped and published all the same, so that no wait
            // behind it lasts (the host reports *err)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            pub_si = si;
            pub_val = k + 1u;
        }
    }
}

// Zeroes the work counter and the n per-slot progress words of one launch.
__global__ void k_reset_queue(uint32_t *queue, uint32_t *done, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) *queue = 0u;
    if (i < n) done[i] = 0u;
}

hipError_t launch_packed_init(PackedLattice *lats, const PackedLattice *stage, const int32_t *idx, int n,
                              hipStream_t s)
{
    // every lattice in the list is initialised with its own L; the grid covers L = 512
    const uint32_t max_words = 2u * cb_words((uint32_t)kResidentMaxL);
    hipLaunchKernelGGL(k_packed_init, dim3((max_words + 255u) / 256u, (uint32_t)n), dim3(256), 0, s,
                       lats, stage, idx);
    return hipGetLastError();
}

template <int NT>
static hipError_t launch_nt(const Slot *slots, uint32_t n_slots, uint32_t n_sub, const int32_t *slot_lats,
                            PackedLattice *lats, uint32_t *queue, uint32_t *done, uint32_t *err,
                            uint32_t lds_bytes, uint32_t grid, hipStream_t s)
{
    // Above 64 KiB of dynamic LDS the kernel has to be opted in.  Batches and
    // servers may launch from different threads: the level only ever rises,
    // under a lock, so a launch never finds it lowered by another.
    static std::mutex mu;
    static uint32_t opted_in = 65536u;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (lds_bytes > opted_in) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_resident<NT>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
            if (e != hipSuccess) return e;
            opted_in = lds_bytes;
        }
    }
    const uint32_t n_done = n_sub > 1u ? n_slots : 0u;
    hipLaunchKernelGGL(k_reset_queue, dim3(n_done / 256u + 1u), dim3(256), 0, s, queue, done, n_done);
    hipLaunchKernelGGL(k_resident<NT>, dim3(grid), dim3(NT), lds_bytes, s, slots, n_slots, n_sub, slot_lats,
                       lats, queue, done, err);
    return hipGetLastError();
}

hipError_t launch_resident(int threads, const Slot *slots, uint32_t n_slots, uint32_t n_sub,
                           const int32_t *slot_lats, PackedLattice *lats, uint32_t *queue, uint32_t *done,
                           uint32_t *err, uint32_t lds_bytes, uint32_t grid, hipStream_t s)
{
    if (n_sub == 0u || (uint64_t)n_slots * n_sub > 0xFFFFFFFFull) return hipErrorInvalidValue;
    switch (threads) {
    case 64: return launch_nt<64>(slots, n_slots, n_sub, slot_lats, lats, queue, done, err, lds_bytes, grid, s);
    case 128: return launch_nt<128>(slots, n_slots, n_sub, slot_lats, lats, queue, done, err, lds_bytes, grid, s);
    case 256: return launch_nt<256>(slots, n_slots, n_sub, slot_lats, lats, queue, done, err, lds_bytes, grid, s);
    case 512: return launch_nt<512>(slots, n_slots, n_sub, slot_lats, lats, queue, done, err, lds_bytes, grid, s);
    case 1024:
        return launch_nt<1024>(slots, n_slots, n_sub, slot_lats, lats, queue, done, err, lds_bytes, grid, s);
    default: return hipErrorInvalidValue;
    }
}

int resident_max_active_groups(int threads, uint32_t lds_bytes)
{
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (threads) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resident<64>, 64, lds_bytes); break;
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resident<128>, 128, lds_bytes); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resident<256>, 256, lds_bytes); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resident<512>, 512, lds_bytes); break;
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resident<1024>, 1024, lds_bytes); break;
    default: break;
    }
    if (e != hipSuccess || per_cu <= 0) return 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace mcrg
