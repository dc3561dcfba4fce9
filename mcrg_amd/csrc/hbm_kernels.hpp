// Descriptor and record layout shared by the host runtime and the HBM-schedule
// kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mcrg {

constexpr uint32_t kMaxLevels = 15;
// Blocked levels of side >= kGridLevelMinL are measured and blocked by a grid
// of workgroups (k_rm_level, one launch per level); the smaller ones by one
// workgroup per lattice (k_levels_accumulate), whose first level (side <= 256,
// 2048 words) is read from global memory and the rest kept in LDS.
constexpr uint32_t kGridLevelMinL = 512;
constexpr uint32_t kWgLevelMaxL = kGridLevelMinL / 2;
// Counters per lattice: level-0 dNN, dNNN, dP, ndown, then dNN, dNNN, dP of
// every grid-measured level l >= 1 at 4 + 3 (l - 1).
constexpr uint32_t kSumWords = 4u + 3u * kMaxLevels;

// Rows per thread strip of the int8 half-sweep (L >= 64).  Up to L = 8192 the
// lattice (<= 64 MiB) stays in the 256 MB last-level cache, so each row's three
// reads (as a row and as the neighbour of two others) cost little and one row
// per thread keeps the most threads in flight.  The 256 MiB of L = 16384 do
// not fit: strips of 8 rows read each row 10/8 times from HBM (131 against
// 159 us per half-sweep; 16 rows: 131, 4: 136, 2: 143).
__host__ __device__ inline uint32_t half_sweep_rows(uint32_t L) { return L >= 16384u ? 8u : 1u; }

// Words of a bit-packed level of side Ll (row-major bits).
__host__ __device__ constexpr uint32_t level_words(uint32_t Ll) { return (Ll * Ll + 31u) / 32u; }
// Word offset of level l >= 1 of a side-L lattice in its level buffer.
__host__ __device__ inline uint32_t level_offset(uint32_t L, uint32_t l)
{
    uint32_t o = 0;
    for (uint32_t j = 1; j < l; ++j) o += level_words(L >> j);
    return o;
}
// Levels 1..n_grid_levels(L, nl) are grid-measured.
__host__ __device__ inline uint32_t n_grid_levels(uint32_t L, uint32_t nl)
{
    uint32_t g = 0;
    while (g < nl && (L >> (g + 1u)) >= kGridLevelMinL) ++g;
    return g;
}

// Per-bin record of one chain measured on levels 0..nl (the reference keeps the
// same sums as FP64, src/mcrg.cpp:59-66, with 3x3 blocks flattened column-major
// as src/definitions.cpp:9-19):
//   S[l][a]                 sum of operator a at level l           3 (nl+1)
//   SELF[l][a][b]           sum of S[l][a] S[l][b]                 9 (nl+1)
//   CROSS[l][a][b], l >= 1  sum of S[l][a] S[l-1][b]               9 nl
//   NN0[l], l >= 1          sum of NN[l] NN[0] (the critical-point  nl
//                           locator's input, src/mcrg.cpp:262-263)
//   M, |M|, M^2 at level 0, sample count                          4
__host__ __device__ constexpr uint32_t off_S(uint32_t l, uint32_t a) { return l * 3u + a; }
__host__ __device__ constexpr uint32_t off_self(uint32_t nl, uint32_t l, uint32_t a, uint32_t b)
{
    return 3u * (nl + 1u) + l * 9u + b * 3u + a;
}
__host__ __device__ constexpr uint32_t off_cross(uint32_t nl, uint32_t l, uint32_t a, uint32_t b)
{
    return 12u * (nl + 1u) + (l - 1u) * 9u + b * 3u + a;
}
__host__ __device__ constexpr uint32_t off_nn0(uint32_t nl, uint32_t l) { return 12u * (nl + 1u) + 9u * nl + l - 1u; }
__host__ __device__ constexpr uint32_t off_M(uint32_t nl) { return 12u * (nl + 1u) + 10u * nl; }
__host__ __device__ constexpr uint32_t record_size(uint32_t nl) { return off_M(nl) + 4u; }

// Entry j of the per-sample contribution to a bin record (layout of off_S,
// off_self, off_cross, off_nn0, off_M) from the operator values S[l][a] and
// the level-0 magnetisation M.
__device__ __forceinline__ int64_t record_term(uint32_t j, uint32_t nl, const int64_t (*S)[3], int64_t M)
{
    const uint32_t n_s = 3u * (nl + 1u), n_self = 9u * (nl + 1u), n_cross = 9u * nl;
    if (j < n_s) return S[j / 3u][j % 3u];
    j -= n_s;
    if (j < n_self) {
        const uint32_t l = j / 9u, a = j % 3u, b = (j % 9u) / 3u;
        return S[l][a] * S[l][b];
    }
    j -= n_self;
    if (j < n_cross) {
        const uint32_t l = j / 9u + 1u, a = j % 3u, b = (j % 9u) / 3u;
        return S[l][a] * S[l - 1u][b];
    }
    j -= n_cross;
    if (j < nl) return S[j + 1u][0] * S[0][0];
    j -= nl;
    return j == 0u ? M : j == 1u ? (M < 0 ? -M : M) : j == 2u ? M * M : 1;
}

struct HbmLattice {
    int8_t *spins;              // L*L int8, row-major (x fastest)
    uint32_t *lvl;              // bit-packed levels 1.. at level_offset(L, l)
    unsigned long long *sums;   // kSumWords disagreement counters
    int64_t *records;           // max_bins * record_size(n_levels)
    uint32_t *state;            // [0] sweep counter t, [1] measured-sample counter
    uint32_t k0, k1;            // Philox key = seed
    Accept acc;
    int L;
    int n_levels;
    int sweeps_per_bin;
    int sampler;                // MCRG_SAMPLER_METROPOLIS or MCRG_SAMPLER_SWENDSEN_WANG
    int32_t *labels;            // L*L cluster labels (Swendsen-Wang only)
    uint32_t bond_thr;          // floor((1 - exp(-2J)) 2^32): bond probability of an aligned pair
    uint32_t pad;
};

// Launchers (host).  `d` points to n_lats device descriptors that all share side L.
hipError_t launch_int8_init(const HbmLattice *d, int n_lats, int L, hipStream_t s);
// One Swendsen-Wang sweep: label reset, bond draw + union, root coin + flip.
hipError_t launch_sw_sweep(const HbmLattice *d, int n_lats, int L, uint32_t t_off, hipStream_t s);
hipError_t launch_int8_half_sweep(const HbmLattice *d, int n_lats, int L, uint32_t t_off,
                                  uint32_t colour, hipStream_t s);
hipError_t launch_int8_measure0(const HbmLattice *d, int n_lats, int L, uint32_t t_off,
                                hipStream_t s);
// Levels 1..nl (grid launches for the large ones, then one workgroup per
// lattice) and the bin record update.
hipError_t launch_levels_accumulate(const HbmLattice *d, int n_lats, int L, int n_levels, uint32_t t_off,
                                    uint32_t q_off, hipStream_t s);
hipError_t launch_advance(const HbmLattice *d, int n_lats, uint32_t sweeps, uint32_t samples,
                          hipStream_t s);

}  // namespace mcrg
