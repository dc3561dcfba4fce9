// LDS-resident schedule: bit-packed (CB) lattices held in LDS by a workgroup
// for a whole chunk of sweeps.  A "slot" is G lattices of one side L and one
// level count that a workgroup advances together; a persistent grid pulls
// slots from a queue.  Shared by the host engine and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "hbm_kernels.hpp"

namespace mcrg {

constexpr int kResidentMinL = 8;
constexpr int kResidentMaxL = 512;
constexpr uint32_t kMetaWords = 16;
constexpr uint32_t kMaxLdsBytes = 160u * 1024u - 64u;

// Per-lattice (job) state in device memory.
struct PackedLattice {
    uint32_t *spins;      // CB words: colour 0 [W] then colour 1 [W], W = L*L/64
    int64_t *records;     // max_bins * record_size(n_levels)
    uint32_t k0, k1;      // Philox key = seed
    Accept acc;
    uint32_t t;           // next sweep index
    uint32_t n_measured;  // measured samples so far
    uint32_t therm_left;  // unmeasured sweeps still to run
    uint32_t meas_left;   // measured sweeps still to run
    int32_t L, n_levels, sweeps_per_bin, max_bins;
};

// G lattice positions of one (L, n_levels) class: slot_lats[first .. first+G),
// -1 marks an empty position.
struct Slot {
    int32_t L, G, n_levels, first;
    uint32_t n_sweeps;  // sweeps this launch advances each lattice of the slot (at most)
};

__host__ __device__ inline uint32_t cb_words(uint32_t L) { return L * L / 64u; }

__host__ __device__ inline uint32_t level_words(uint32_t L, uint32_t l)
{
    const uint32_t Ll = L >> l;
    const uint32_t w = (Ll * Ll + 31u) / 32u;
    return w ? w : 1u;
}

// LDS layout of one slot, offsets in 32-bit words.  T = threads per lattice.
struct SlotLayout {
    uint32_t W;           // CB words per colour
    uint32_t lvl_words;   // RM words of levels 1..nl, per lattice (also the coin words)
    uint32_t team_waves;  // waves per lattice team (count partials per level)
    uint32_t n_sv;        // operator values per lattice: 3 (nl+1) + M, |M|, 1
    uint32_t coin_calls;  // Philox calls for all tie-coin words (4 words per call)
    uint32_t meta, spins, lvl, coin, cnt, sv, pairs, rec, total;
};

__host__ __device__ inline SlotLayout slot_layout(uint32_t L, uint32_t G, uint32_t nl, uint32_t T)
{
    SlotLayout s;
    s.W = cb_words(L);
    s.lvl_words = 0;
    s.coin_calls = 0;
    for (uint32_t l = 1; l <= nl; ++l) {
        s.lvl_words += level_words(L, l);
        s.coin_calls += (level_words(L, l) + 3u) / 4u;
    }
    s.team_waves = T > 64u ? T / 64u : 1u;
    s.meta = 0;
    s.spins = s.meta + G * kMetaWords;
    s.lvl = s.spins + G * 2u * s.W;
    s.coin = s.lvl + G * s.lvl_words;
    s.cnt = s.coin + G * s.lvl_words;
    s.n_sv = 3u * (nl + 1u) + 3u;
    s.sv = s.cnt + G * s.team_waves * (nl + 1u) * 4u;  // int32 operator values (teams within one wave)
    s.pairs = s.sv + G * s.n_sv;                         // one (x, y) operand pair per record entry
    s.rec = (s.pairs + record_size(nl) + 1u) & ~1u;  // int64 aligned
    s.total = s.rec + G * record_size(nl) * 2u;
    return s;
}

// First blocked level whose rows fit one per lane of a team of T lanes (at
// most one wave), or 0 when no level from 1 to nl does: the levels from it on
// are measured in registers (the "pyramid").
__host__ __device__ inline uint32_t pyramid_level(uint32_t L, uint32_t T, uint32_t nl)
{
    const uint32_t lanes = T < 64u ? T : 64u;
    for (uint32_t l = 1; l <= nl; ++l)
        if ((L >> l) <= lanes) return l;
    return 0u;
}

// Slot bodies the resident kernel is compiled with: the wide body for teams of
// at least a wave on L >= 64, whose register pyramid starts at rows of 64
// (L >= 128) or 32 sites (L = 64), unrolled for those two shapes and without
// the small-lattice code paths; the generic body for everything else.
enum SlotShape : uint32_t { kShapeGeneric = 0, kShapeRows64 = 1, kShapeRows32 = 2 };

__host__ __device__ inline uint32_t slot_shape(uint32_t L, uint32_t G, uint32_t nl, uint32_t NT)
{
    const uint32_t T = NT / G;
    const uint32_t pl = pyramid_level(L, T, nl);
    if (pl == 0u || T < 64u) return kShapeGeneric;
    const uint32_t Lp = L >> pl;
    return Lp == 64u ? kShapeRows64 : Lp == 32u && L >= 64u ? kShapeRows32 : kShapeGeneric;
}

// Admits n lattices: lats[idx[k]] = stage[k], its records zeroed, its spins
// drawn (stage and idx in device memory).
hipError_t launch_packed_init(PackedLattice *lats, const PackedLattice *stage, const int32_t *idx, int n,
                              hipStream_t s);
// Work items of one launch: sub-chunk k (0 <= k < n_sub) of slot i is item
// k * n_slots + i, so that every slot's first part is queued before any second
// part.  Sub-chunk k runs sweeps [k c, min((k+1) c, n_sweeps)) of its slot,
// c = ceil(n_sweeps / n_sub), and starts only after sub-chunk k-1 of the same
// slot has published its write-back (done[i] = k); with n_sub = 1 there is no
// waiting.  `done` holds n_slots words and `queue` one, both zeroed by the
// launch; a wait that exceeds kSubChunkWaitTicks sets *err.
constexpr uint64_t kSubChunkWaitTicks = 2000000000ull;  // 20 s of the 100 MHz clock
hipError_t launch_resident(int threads, const Slot *slots, uint32_t n_slots, uint32_t n_sub,
                           const int32_t *slot_lats, PackedLattice *lats, uint32_t *queue, uint32_t *done,
                           uint32_t *err, uint32_t lds_bytes, uint32_t grid, hipStream_t s);
int resident_max_active_groups(int threads, uint32_t lds_bytes);

}  // namespace mcrg
