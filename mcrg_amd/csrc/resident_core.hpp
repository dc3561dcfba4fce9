// Body of the LDS-resident schedule: one workgroup advances the G lattices of
// one slot by up to the slot's n_sweeps sweeps, measuring the sweeps that fall in each
// lattice's measurement phase.
//
// Each lattice is owned by a team of T = nt / G consecutive threads (lane u of
// lattice g is thread g*T + u).  A team of at most 64 lanes sits inside one
// wave and synchronises at wave level only; larger teams (L >= 128) use the
// workgroup barrier.  Per measured sweep:
//   coins    every tie-coin word of every blocking level, one Philox call per
//            4 consecutive coin words (the coins do not depend on the spins)
//   level 0  operator counts of the CB lattice, blocking into the RM level 1
//   levels   operator counts and blocking of RM levels 1..nl
//   record   team-wide sums of the per-lane counts, then one lane per record
//            entry adds the sample's sums and products to the LDS partial
//            record, which is flushed into the per-bin record in HBM when the
//            lattice's sample index enters a new bin and at the chunk end.
//
// Loops are written as `for g` / `for u` with strides (gs, us); each runs once
// per thread (g = tid / T, u = tid % T).
#pragma once

#include "levels.hpp"
#include "resident.hpp"

namespace mcrg {

// Meta words per lattice position (LDS).
enum : uint32_t {
    kMetaIdx = 0,     // lattice index, 0xFFFFFFFF for an empty position
    kMetaT = 1,       // sweep index at chunk start
    kMetaTherm = 2,   // unmeasured sweeps left at chunk start
    kMetaMeas = 3,    // measured sweeps left at chunk start
    kMetaNMeas = 4,   // measured samples before the chunk
    kMetaSpb = 5,     // sweeps per bin
    kMetaK0 = 6,
    kMetaK1 = 7,
    kMetaThr0 = 8,
    kMetaThr1 = 9,
    kMetaAlw0 = 10,
    kMetaAlw1 = 11,
};

struct Team {
    uint32_t g0, gs, u0, us;  // loop starts and strides over lattices and lanes
    uint32_t T;               // lanes per lattice
    uint32_t lane;            // thread index within the workgroup
    bool block;               // team spans more than one wave
};

__device__ __forceinline__ void team_sync(const Team &tm)
{
    if (tm.block) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__device__ __forceinline__ bool team_any(const Team &tm, bool p)
{
    if (tm.block) return __syncthreads_or(p ? 1 : 0) != 0;
    return __ballot(p ? 1 : 0) != 0ull;
}

__device__ __forceinline__ void block_sync() { __syncthreads(); }

// Sum over the aligned group of `width` lanes (1..64, a power of two) that
// contains this lane; for width 64 the total is valid in lanes 0..31.
__device__ __forceinline__ uint32_t group_sum(uint32_t v, uint32_t width)
{
    if (width >= 2u) v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    if (width >= 4u) v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    if (width >= 8u) v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false); // row_half_mirror
    if (width >= 16u) v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false); // row_mirror
    if (width >= 32u) v += (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);             // xor 16
    if (width >= 64u) v += (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    return v;
}

// The unmeasured and measured sweeps left (kMetaTherm, kMetaMeas: adjacent
// words, 8-byte aligned in each lattice's 64-byte meta block) in one LDS read.
static_assert(kMetaMeas == kMetaTherm + 1u && kMetaTherm % 2u == 0u && kMetaWords % 2u == 0u,
              "sweeps_left reads kMetaTherm and kMetaMeas as one aligned 64-bit word");
__device__ __forceinline__ uint2 sweeps_left(const uint32_t *m)
{
    return *reinterpret_cast<const uint2 *>(m + kMetaTherm);
}
__device__ __forceinline__ bool in_run(const uint32_t *m, uint32_t it)
{
    const uint2 v = sweeps_left(m);
    return it < v.x + v.y;
}
__device__ __forceinline__ bool in_meas(const uint32_t *m, uint32_t it)
{
    const uint2 v = sweeps_left(m);
    return it >= v.x && it < v.x + v.y;
}

// Operator counts of lattice g at level l: sum over the team's wave partials;
// levels >= lf were measured by the team's first wave alone.
__device__ __forceinline__ uint32_t count_of(const uint32_t *cntg, uint32_t nw, uint32_t nl, uint32_t l,
                                             uint32_t a, uint32_t lf)
{
    if (l >= lf) nw = 1u;
    uint32_t s = 0;
    for (uint32_t w = 0; w < nw; ++w) s += cntg[(w * (nl + 1u) + l) * 4u + a];
    return s;
}

__device__ __forceinline__ int64_t level_S(const uint32_t *cntg, uint32_t nw, uint32_t nl, uint32_t L,
                                           uint32_t l, uint32_t a, uint32_t lf)
{
    const int64_t n = (int64_t)(L >> l) * (int64_t)(L >> l);
    const int64_t d = (int64_t)count_of(cntg, nw, nl, l, a, lf);
    return a < 2u ? 2 * n - 2 * d : n - 2 * d;
}

// Record entry k of one sample is sv[x] * sv[y] over the lattice's operator
// values sv = (S[0][0..2], ..., S[nl][0..2], M, |M|, 1) (layout of record_size():
// S, SELF, CROSS, NN0, then M, |M|, M^2 and the count).
__device__ __forceinline__ void record_pair(uint32_t nl, uint32_t k, uint32_t &x, uint32_t &y)
{
    const uint32_t n_s = 3u * (nl + 1u), n_self = 9u * (nl + 1u), n_cross = 9u * nl, n_nn0 = nl;
    const uint32_t iM = n_s, iAbs = n_s + 1u, iOne = n_s + 2u;
    if (k < n_s) {
        x = k;
        y = iOne;
    } else if (k < n_s + n_self) {
        const uint32_t kk = k - n_s, l = kk / 9u, r = kk % 9u;
        x = 3u * l + r % 3u;
        y = 3u * l + r / 3u;
    } else if (k < n_s + n_self + n_cross) {
        const uint32_t kk = k - n_s - n_self, l = kk / 9u + 1u, r = kk % 9u;
        x = 3u * l + r % 3u;
        y = 3u * (l - 1u) + r / 3u;
    } else if (k < n_s + n_self + n_cross + n_nn0) {
        x = 3u * (k - n_s - n_self - n_cross + 1u);  // NN of level l
        y = 0u;                                       // NN of level 0
    } else {
        const uint32_t m = k - n_s - n_self - n_cross - n_nn0;
        x = m == 0u ? iM : m == 1u ? iAbs : m == 2u ? iM : iOne;
        y = m == 2u ? iM : iOne;
    }
}

// Operator value i of sv (see record_pair) from the counts.  |value| <= 2 L^2
// <= 2^19 for L <= 512, so the slot keeps them as int32.
// Levels >= lf come from the first wave's partial only (the register pyramid);
// lf = nl + 1 sums every level over all partials.
__device__ __forceinline__ int64_t sv_value(const uint32_t *cntg, uint32_t nw, uint32_t nl, uint32_t L, uint32_t i,
                                            uint32_t lf)
{
    const uint32_t n_s = 3u * (nl + 1u);
    if (i < n_s) return level_S(cntg, nw, nl, L, i / 3u, i % 3u, lf);
    if (i == n_s + 2u) return 1;
    const int64_t n0 = (int64_t)L * (int64_t)L;
    const int64_t M = n0 - 2 * (int64_t)count_of(cntg, nw, nl, 0u, 3u, lf);
    return i == n_s ? M : (M < 0 ? -M : M);
}

// Team-wide reduction of one level's counts into the lattice's wave partial.
__device__ __forceinline__ void store_counts(uint32_t *cntg, const Team &tm, uint32_t nl, uint32_t l,
                                             uint32_t u, uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3)
{
    const uint32_t width = tm.T < 64u ? tm.T : 64u;
    d0 = group_sum(d0, width);
    d1 = group_sum(d1, width);
    d2 = group_sum(d2, width);
    d3 = group_sum(d3, width);
    if ((u & (width - 1u)) != 0u) return;
    uint32_t *c = cntg + ((u / width) * (nl + 1u) + l) * 4u;
    c[0] = d0;
    c[1] = d1;
    c[2] = d2;
    c[3] = d3;
}

// Wave-uniform value: when a lattice's team covers whole waves its key and
// thresholds are the same in every lane and can live in scalar registers.
template <bool UNIFORM>
__device__ __forceinline__ uint32_t uni(uint32_t v)
{
    return UNIFORM ? (uint32_t)__builtin_amdgcn_readfirstlane((int)v) : v;
}

// Half-sweep of colour c for the team's lattice.
template <bool UNIFORM>
__device__ __forceinline__ void sweep_colour(const uint32_t *meta, uint32_t *spins, uint32_t L, uint32_t W,
                                             uint32_t G, const Team &tm, uint32_t it, uint32_t c)
{
    for (uint32_t g = tm.g0; g < G; g += tm.gs) {
        const uint32_t *m = meta + g * kMetaWords;
        if (!in_run(m, it)) continue;
        uint32_t *own = spins + g * 2u * W + c * W;
        const uint32_t *oth = spins + g * 2u * W + (1u - c) * W;
        Accept acc;
        acc.thr0 = uni<UNIFORM>(m[kMetaThr0]);
        acc.thr1 = uni<UNIFORM>(m[kMetaThr1]);
        acc.always0 = uni<UNIFORM>(m[kMetaAlw0]);
        acc.always1 = uni<UNIFORM>(m[kMetaAlw1]);
        const uint32_t k0 = uni<UNIFORM>(m[kMetaK0]), k1 = uni<UNIFORM>(m[kMetaK1]);
        const uint32_t t = uni<UNIFORM>(m[kMetaT]) + it;
        for (uint32_t u = tm.u0; u < tm.T; u += tm.us) {
            for (uint32_t w = u; w < W; w += tm.T) {
                const Nbr4 n = cb_neighbours(oth, L, W, w, c);
                own[w] ^= metropolis_word(own[w], n.nu, n.nd, n.nl, n.nr, 0xFFFFFFFFu, acc, k0, k1, w, t, c);
            }
        }
    }
}

// Row y+1 (mod Ll) for lane y.  A team of a whole wave rotates the wave by one
// lane with DPP (no LDS round trip) and patches the wrap-around row; smaller
// teams shuffle within their segment.
__device__ __forceinline__ uint32_t row_next(uint32_t v, uint32_t y, uint32_t Ll, uint32_t width)
{
    if (width == 64u) {
        const uint32_t n = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x134, 0xF, 0xF, false);  // wave_rol:1
        const uint32_t first = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
        return y == Ll - 1u ? first : n;
    }
    return (uint32_t)__shfl((int)v, (int)((y + 1u) & (Ll - 1u)), (int)width);
}

// Levels pl..nl of one lattice in registers: lane y (< 64, within the team's
// first wave) holds row y of the current level (rows of 64 bits in two words).
// The row below comes from row_next; lane 2by forms blocked row by from its own
// row and the row below, and one shuffle moves it to lane by.  Per level the
// NN and NNN counts travel packed in one register (each < 2^14 for rows of at
// most 64 sites) and are summed with DPP.  `width` is the shuffle segment (the
// team, or 64 when the team spans waves).
__device__ __forceinline__ void pyramid(const uint32_t *lg, const uint32_t *cg, uint32_t *cntg, uint32_t L,
                                        uint32_t pl, uint32_t nl, uint32_t lvl_off, uint32_t coin_off,
                                        uint32_t y, uint32_t width)
{
    uint32_t Ll = L >> pl;
    uint32_t lo = 0, hi = 0;
    if (y < Ll) {
        if (Ll == 64u) {
            lo = lg[lvl_off + 2u * y];
            hi = lg[lvl_off + 2u * y + 1u];
        } else {
            lo = rm_row_small(lg + lvl_off, Ll, y);
        }
    }
    uint32_t coff = coin_off;
    for (uint32_t l = pl; l <= nl; ++l) {
        const bool wide = Ll == 64u;
        const uint32_t Lb = Ll >> 1;
        // coin word of the blocks this lane forms (lane 2by forms blocked row by)
        uint32_t coin = 0;
        if (l < nl && y < Ll && (y & 1u) == 0u) {
            const uint32_t bm = (y >> 1) * Lb;
            coin = cg[coff + (bm >> 5)] >> (bm & 31u);
        }
        // ---- operators of level l
        const uint32_t dlo = row_next(lo, y, Ll, width);
        const uint32_t dhi = wide ? row_next(hi, y, Ll, width) : 0u;
        uint32_t nn_nnn = 0, dp = 0;
        if (y < Ll) {
            if (wide) {
                const uint32_t rlo = (lo >> 1) | (hi << 31), rhi = (hi >> 1) | (lo << 31);
                const uint32_t glo = (dlo >> 1) | (dhi << 31), ghi = (dhi >> 1) | (dlo << 31);
                const uint32_t dnn = (uint32_t)(popc(lo ^ rlo) + popc(hi ^ rhi) + popc(lo ^ dlo) + popc(hi ^ dhi));
                const uint32_t dnnn =
                    (uint32_t)(popc(lo ^ glo) + popc(hi ^ ghi) + popc(rlo ^ dlo) + popc(rhi ^ dhi));
                nn_nnn = dnn | (dnnn << 16);
                dp = (uint32_t)(popc(xor3(lo, rlo, dlo) ^ glo) + popc(xor3(hi, rhi, dhi) ^ ghi));
            } else {
                const uint32_t r = Ll == 32u ? ((lo >> 1) | (lo << 31)) : rot_right_small(lo, Ll);
                const uint32_t g = Ll == 32u ? ((dlo >> 1) | (dlo << 31)) : rot_right_small(dlo, Ll);
                nn_nnn = (uint32_t)(popc(lo ^ r) + popc(lo ^ dlo)) | ((uint32_t)(popc(lo ^ g) + popc(r ^ dlo)) << 16);
                dp = (uint32_t)popc(xor3(lo, r, dlo) ^ g);
            }
        }
        // rows y >= Ll contribute nothing: sum over the level's Ll lanes
        nn_nnn = group_sum(nn_nnn, Ll);
        dp = group_sum(dp, Ll);
        if (y == 0u) {
            uint32_t *c = cntg + l * 4u;
            c[0] = nn_nnn & 0xFFFFu;
            c[1] = nn_nnn >> 16;
            c[2] = dp;
            c[3] = 0u;
        }
        if (l == nl) break;
        // ---- blocking into level l + 1
        uint32_t blk = 0;
        if (y < Ll && (y & 1u) == 0u) {
            if (wide)
                blk = block16(lo, dlo, coin & 0xFFFFu) | (block16(hi, dhi, coin >> 16) << 16);
            else
                blk = block16(lo, dlo, coin) & ((1u << Lb) - 1u);
        }
        const uint32_t nlo = (uint32_t)__shfl((int)blk, (int)((2u * y) & (width - 1u)), (int)width);
        coff += level_words(L, l + 1u);
        lo = y < Lb ? nlo : 0u;
        hi = 0u;
        Ll = Lb;
    }
}

// pyramid() for a starting row width LP and shuffle segment WIDTH known at
// compile time: its levels are unrolled, so every width test (the group sums,
// the row rotation, the 64-site rows) folds away.  The same operations in the
// same order, so the counts and blocked rows are identical.
template <uint32_t LP, uint32_t WIDTH>
__device__ __forceinline__ void pyramid_t(const uint32_t *lg, const uint32_t *cg, uint32_t *cntg, uint32_t L,
                                          uint32_t pl, uint32_t nl, uint32_t lvl_off, uint32_t coin_off,
                                          uint32_t y)
{
    uint32_t lo = 0, hi = 0;
    if (y < LP) {
        if (LP == 64u) {
            lo = lg[lvl_off + 2u * y];
            hi = lg[lvl_off + 2u * y + 1u];
        } else {
            lo = rm_row_small(lg + lvl_off, LP, y);
        }
    }
    uint32_t coff = coin_off;
#pragma unroll
    for (uint32_t j = 0; (LP >> j) >= 2u; ++j) {
        const uint32_t l = pl + j;
        if (l > nl) break;
        const uint32_t Ll = LP >> j;
        const bool wide = Ll == 64u;
        const uint32_t Lb = Ll >> 1;
        // the coin word of the blocks lane y forms (lane 2by forms blocked row
        // by), loaded by every lane (rows y >= Ll wrap into the level) under a
        // wave-uniform test instead of a lane-divergent one
        uint32_t coin = 0;
        if (l < nl) {
            const uint32_t bm = ((y & (Ll - 1u)) >> 1) * Lb;
            coin = cg[coff + (bm >> 5)] >> (bm & 31u);
        }
        const uint32_t dlo = row_next(lo, y, Ll, WIDTH);
        const uint32_t dhi = wide ? row_next(hi, y, Ll, WIDTH) : 0u;
        uint32_t nn_nnn = 0, dp = 0;
        if (y < Ll) {
            if (wide) {
                const uint32_t rlo = (lo >> 1) | (hi << 31), rhi = (hi >> 1) | (lo << 31);
                const uint32_t glo = (dlo >> 1) | (dhi << 31), ghi = (dhi >> 1) | (dlo << 31);
                const uint32_t dnn = (uint32_t)(popc(lo ^ rlo) + popc(hi ^ rhi) + popc(lo ^ dlo) + popc(hi ^ dhi));
                const uint32_t dnnn =
                    (uint32_t)(popc(lo ^ glo) + popc(hi ^ ghi) + popc(rlo ^ dlo) + popc(rhi ^ dhi));
                nn_nnn = dnn | (dnnn << 16);
                dp = (uint32_t)(popc(xor3(lo, rlo, dlo) ^ glo) + popc(xor3(hi, rhi, dhi) ^ ghi));
            } else {
                const uint32_t r = Ll == 32u ? ((lo >> 1) | (lo << 31)) : rot_right_small(lo, Ll);
                const uint32_t g = Ll == 32u ? ((dlo >> 1) | (dlo << 31)) : rot_right_small(dlo, Ll);
                nn_nnn = (uint32_t)(popc(lo ^ r) + popc(lo ^ dlo)) | ((uint32_t)(popc(lo ^ g) + popc(r ^ dlo)) << 16);
                dp = (uint32_t)popc(xor3(lo, r, dlo) ^ g);
            }
        }
        nn_nnn = group_sum(nn_nnn, Ll);
        dp = group_sum(dp, Ll);
        if (y == 0u) {
            uint32_t *c = cntg + l * 4u;
            c[0] = nn_nnn & 0xFFFFu;
            c[1] = nn_nnn >> 16;
            c[2] = dp;
            c[3] = 0u;
        }
        if (l == nl) break;
        uint32_t blk = 0;
        if (y < Ll && (y & 1u) == 0u) {
            if (wide)
                blk = block16(lo, dlo, coin & 0xFFFFu) | (block16(hi, dhi, coin >> 16) << 16);
            else
                blk = block16(lo, dlo, coin) & ((1u << Lb) - 1u);
        }
        const uint32_t nlo = (uint32_t)__shfl((int)blk, (int)((2u * y) & (WIDTH - 1u)), (int)WIDTH);
        coff += level_words(L, l + 1u);
        lo = y < Lb ? nlo : 0u;
        hi = 0u;
    }
}

template <uint32_t SHAPE>
__device__ void run_slot(const Slot sl, const int32_t *__restrict__ slot_lats,
                         PackedLattice *__restrict__ lats, uint32_t *lds, Team tm)
{
    const uint32_t L = (uint32_t)sl.L, G = (uint32_t)sl.G, nl = (uint32_t)sl.n_levels;
    const uint32_t n_chunk = sl.n_sweeps;
    const uint32_t T = tm.T;
    if constexpr (SHAPE != kShapeGeneric) {
        // what the shape implies (slot_shape): teams of at least a wave and
        // L >= 64, so the kernel keeps only those code paths
        __builtin_assume(T >= 64u);
        __builtin_assume(L >= 64u);
    }
    const SlotLayout lay = slot_layout(L, G, nl, T);
    const uint32_t W = lay.W, rs = record_size(nl), nw = lay.team_waves;
    const uint32_t cstride = nw * (nl + 1u) * 4u;
    uint32_t *meta = lds + lay.meta;
    uint32_t *spins = lds + lay.spins;
    uint32_t *lvl = lds + lay.lvl;
    uint32_t *coins = lds + lay.coin;
    uint32_t *cnt = lds + lay.cnt;
    int32_t *sv = reinterpret_cast<int32_t *>(lds + lay.sv);
    int64_t *rec = reinterpret_cast<int64_t *>(lds + lay.rec);
    uint32_t *pairs = lds + lay.pairs;

    // ---- load: meta, spins, zero partial records
    for (uint32_t g = tm.g0; g < G; g += tm.gs) {
        for (uint32_t u = tm.u0; u < 1u; u += tm.us) {
            uint32_t *m = meta + g * kMetaWords;
            const int32_t li = slot_lats[sl.first + (int32_t)g];
            if (li >= 0) {
                const PackedLattice &P = lats[li];
                m[kMetaIdx] = (uint32_t)li;
                m[kMetaT] = P.t;
                m[kMetaTherm] = P.therm_left;
                m[kMetaMeas] = P.meas_left;
                m[kMetaNMeas] = P.n_measured;
                m[kMetaSpb] = (uint32_t)P.sweeps_per_bin;
                m[kMetaK0] = P.k0;
                m[kMetaK1] = P.k1;
                m[kMetaThr0] = P.acc.thr0;
                m[kMetaThr1] = P.acc.thr1;
                m[kMetaAlw0] = P.acc.always0;
                m[kMetaAlw1] = P.acc.always1;
            } else {
                m[kMetaIdx] = 0xFFFFFFFFu;
                m[kMetaTherm] = 0u;
                m[kMetaMeas] = 0u;
            }
        }
    }
    block_sync();
    for (uint32_t g = tm.g0; g < G; g += tm.gs) {
        const uint32_t li = meta[g * kMetaWords + kMetaIdx];
        for (uint32_t u = tm.u0; u < T; u += tm.us) {
            for (uint32_t j = u; j < 2u * W; j += T)
                spins[g * 2u * W + j] = li != 0xFFFFFFFFu ? lats[li].spins[j] : 0u;
            for (uint32_t k = u; k < rs; k += T) rec[g * rs + k] = 0;
            for (uint32_t k = g * T + u; k < rs; k += G * T) {
                uint32_t x, y;
                record_pair(nl, k, x, y);
                pairs[k] = x | (y << 16);
            }
        }
    }
    block_sync();

    for (uint32_t it = 0; it < n_chunk; ++it) {
        // Whether any lattice still runs / measures this sweep.  A block-wide
        // team (one or two lattices per workgroup) evaluates it for every
        // lattice of the slot in every thread, so the answer is uniform without
        // a barrier; wave-level teams combine their lanes with a ballot.
        bool running = false, measuring = false;
        if (tm.block) {
            for (uint32_t g = 0; g < G; ++g) {
                running = running || in_run(meta + g * kMetaWords, it);
                measuring = measuring || in_meas(meta + g * kMetaWords, it);
            }
        } else {
            for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                running = running || in_run(meta + g * kMetaWords, it);
                measuring = measuring || in_meas(meta + g * kMetaWords, it);
            }
            running = team_any(tm, running);
            measuring = team_any(tm, measuring);
        }
        if (!running) break;

        // ---- sweep: colour 0, then colour 1; the tie coins of this sample's
        //      blockings are drawn in the colour-1 phase (they do not depend on
        //      the spins): call c of the flattened list gives coin words
        //      4q..4q+3 of its level
        for (uint32_t c = 0; c < 2u; ++c) {
            if (SHAPE != kShapeGeneric || T >= 64u)
                sweep_colour<true>(meta, spins, L, W, G, tm, it, c);
            else
                sweep_colour<false>(meta, spins, L, W, G, tm, it, c);
            if (c == 1u && measuring) {
                for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                    const uint32_t *m = meta + g * kMetaWords;
                    if (!in_meas(m, it)) continue;
                    const uint32_t t = m[kMetaT] + it;
                    for (uint32_t u = tm.u0; u < T; u += tm.us) {
                        for (uint32_t cc = u; cc < lay.coin_calls; cc += T) {
                            uint32_t l = 0, q = cc, off = 0;
                            for (; l + 1u < nl; ++l) {
                                const uint32_t calls = (level_words(L, l + 1u) + 3u) / 4u;
                                if (q < calls) break;
                                q -= calls;
                                off += level_words(L, l + 1u);
                            }
                            const uint32_t n_words = level_words(L, l + 1u);
                            const u32x4 r =
                                philox(q, t, 0u, (kTagTie << 28) | (l << 16), m[kMetaK0], m[kMetaK1]);
                            uint32_t *cw = coins + g * lay.lvl_words + off + 4u * q;
                            cw[0] = r.v0;
                            if (4u * q + 1u < n_words) cw[1] = r.v1;
                            if (4u * q + 2u < n_words) cw[2] = r.v2;
                            if (4u * q + 3u < n_words) cw[3] = r.v3;
                        }
                    }
                }
            }
            team_sync(tm);
        }
        if (!measuring) continue;

        // The sample's sums and products into lattice g's partial record,
        // entries k = k0, k0 + ks, ...: from the operator values in `sv`, or
        // each entry forming its two values from the counts.
        auto record_sample = [&](uint32_t g, uint32_t k0, uint32_t ks, bool use_sv) {
            const uint32_t *m = meta + g * kMetaWords;
            const uint32_t q = m[kMetaNMeas] + (it - m[kMetaTherm]);
            const uint32_t spb = m[kMetaSpb];
            const bool new_bin = q % spb == 0u && q > m[kMetaNMeas];
            const uint32_t *cntg = cnt + g * cstride;
            for (uint32_t k = k0; k < rs; k += ks) {
                int64_t &r = rec[g * rs + k];
                if (new_bin) {
                    lats[m[kMetaIdx]].records[(size_t)((q - 1u) / spb) * rs + k] += r;
                    r = 0;
                }
                const uint32_t xy = pairs[k];
                if (use_sv)
                    r += (int64_t)sv[g * lay.n_sv + (xy & 0xFFFFu)] * (int64_t)sv[g * lay.n_sv + (xy >> 16)];
                else
                    r += sv_value(cntg, nw, nl, L, xy & 0xFFFFu, nl + 1u) * sv_value(cntg, nw, nl, L, xy >> 16, nl + 1u);
            }
        };

        // ---- level 0: operator counts and blocking into level 1
        for (uint32_t g = tm.g0; g < G; g += tm.gs) {
            const uint32_t *m = meta + g * kMetaWords;
            const bool meas = in_meas(m, it);
            const uint32_t *c0 = spins + g * 2u * W, *c1 = c0 + W;
            for (uint32_t u = tm.u0; u < T; u += tm.us) {
                uint32_t dnn = 0, dnnn = 0, dp = 0, nd = 0;
                if (meas) {
                    for (uint32_t w = u; w < W; w += T) {
                        cb_ops_item(c0, c1, L, W, w, 0u, dnn, dnnn, dp, nd);
                        cb_ops_item(c1, c0, L, W, w, 1u, dnn, dnnn, dp, nd);
                    }
                    if (nl > 0u) {
                        const uint32_t nb = cb_block_items(L);
                        uint32_t *l1 = lvl + g * lay.lvl_words;
                        const uint32_t *cw = coins + g * lay.lvl_words;
                        for (uint32_t j = u; j < nb; j += T) l1[j] = cb_block_item(c0, c1, L, W, j, cw[j]);
                    }
                }
                store_counts(cnt + g * cstride, tm, nl, 0u, u, dnn, dnnn, dp, nd);
            }
        }
        team_sync(tm);

        // ---- levels 1..nl: spread over the team while a level has more rows
        //      than the team's first wave has lanes, then the register pyramid
        const uint32_t pl = pyramid_level(L, T, nl);
        const uint32_t last_spread = pl ? pl - 1u : nl;
        uint32_t off = 0;
        for (uint32_t l = 1; l <= last_spread; ++l) {
            const uint32_t Ll = L >> l;
            const uint32_t nxt = off + level_words(L, l);
            for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                const bool meas = in_meas(meta + g * kMetaWords, it);
                uint32_t *lg = lvl + g * lay.lvl_words;
                const uint32_t *cw = coins + g * lay.lvl_words + nxt;
                for (uint32_t u = tm.u0; u < T; u += tm.us) {
                    uint32_t dnn = 0, dnnn = 0, dp = 0;
                    if (meas) {
                        const uint32_t n_ops = rm_ops_items(Ll);
                        for (uint32_t i = u; i < n_ops; i += T) rm_ops_item(lg + off, Ll, i, dnn, dnnn, dp);
                        if (l < nl) {
                            const uint32_t nb = rm_block_items(Ll);
                            for (uint32_t j = u; j < nb; j += T) lg[nxt + j] = rm_block_item(lg + off, Ll, j, cw[j]);
                        }
                    }
                    store_counts(cnt + g * cstride, tm, nl, l, u, dnn, dnnn, dp, 0u);
                }
            }
            off = nxt;
            team_sync(tm);
        }

        if (pl) {
            // `off` is now the RM offset of level pl; its coins start one level later
            const uint32_t coin_off = off + level_words(L, pl);
            for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                const bool meas = in_meas(meta + g * kMetaWords, it);
                for (uint32_t u = tm.u0; u < T; u += tm.us) {
                    uint32_t *cntg = cnt + g * cstride + (u >> 6) * (nl + 1u) * 4u;
                    if (u < 64u && meas) {
                        if constexpr (SHAPE != kShapeGeneric) {
                            if ((L >> pl) == 64u)
                                pyramid_t<64u, 64u>(lvl + g * lay.lvl_words, coins + g * lay.lvl_words, cntg, L,
                                                    pl, nl, off, coin_off, u);
                            else
                                pyramid_t<32u, 64u>(lvl + g * lay.lvl_words, coins + g * lay.lvl_words, cntg, L,
                                                    pl, nl, off, coin_off, u);
                        } else
                            pyramid(lvl + g * lay.lvl_words, coins + g * lay.lvl_words, cntg, L, pl, nl, off,
                                    coin_off, u, T < 64u ? T : 64u);
                        if (tm.block) {
                            // A team spanning waves: its first wave, which holds
                            // the pyramid's counts (the other partials were
                            // published by earlier barriers), forms the operator
                            // values and updates the record alone.  No barrier
                            // follows: the other waves go on to the next sweep,
                            // which touches neither these buffers nor the
                            // counts before the first wave has met them at the
                            // end of that sweep's first half.
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                            for (uint32_t i = u; i < lay.n_sv; i += 64u)
                                sv[g * lay.n_sv + i] = (int32_t)sv_value(cntg, nw, nl, L, i, pl);
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                            record_sample(g, u, 64u, true);
                        }
                    }
                }
            }
            if (!tm.block) team_sync(tm);
        }
        // a team spanning waves with a pyramid has updated its record above
        const bool recorded = tm.block && pl != 0u;

        // ---- the sample's sums and products into the partial records.  A
        //      team within one wave first forms each operator value once (its
        //      barrier is a wave barrier).  No barrier follows: the next writes
        //      of the counts and values (next sample) come after the two sweep
        //      barriers.
        if (!tm.block) {
            for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                if (!in_meas(meta + g * kMetaWords, it)) continue;
                const uint32_t *cntg = cnt + g * cstride;
                for (uint32_t u = tm.u0; u < T; u += tm.us)
                    for (uint32_t i = u; i < lay.n_sv; i += T)
                        sv[g * lay.n_sv + i] = (int32_t)sv_value(cntg, nw, nl, L, i, nl + 1u);
            }
            team_sync(tm);
        }
        if (!recorded) {
            for (uint32_t g = tm.g0; g < G; g += tm.gs) {
                if (!in_meas(meta + g * kMetaWords, it)) continue;
                for (uint32_t u = tm.u0; u < T; u += tm.us) record_sample(g, u, T, !tm.block);
            }
        }
    }
    block_sync();

    // ---- write back: partial records, spins, counters
    for (uint32_t g = tm.g0; g < G; g += tm.gs) {
        const uint32_t *m = meta + g * kMetaWords;
        if (m[kMetaIdx] == 0xFFFFFFFFu) continue;
        PackedLattice &P = lats[m[kMetaIdx]];
        const uint32_t total = m[kMetaTherm] + m[kMetaMeas];
        const uint32_t swept = total < n_chunk ? total : n_chunk;
        const uint32_t measured = swept > m[kMetaTherm] ? swept - m[kMetaTherm] : 0u;
        for (uint32_t u = tm.u0; u < T; u += tm.us) {
            if (measured > 0u) {
                const uint32_t q_last = m[kMetaNMeas] + measured - 1u;
                for (uint32_t k = u; k < rs; k += T)
                    P.records[(size_t)(q_last / m[kMetaSpb]) * rs + k] += rec[g * rs + k];
            }
            for (uint32_t j = u; j < 2u * W; j += T) P.spins[j] = spins[g * 2u * W + j];
            if (u == 0u) {
                P.t = m[kMetaT] + swept;
                P.n_measured = m[kMetaNMeas] + measured;
                P.therm_left = m[kMetaTherm] - (swept - measured);
                P.meas_left = m[kMetaMeas] - measured;
            }
        }
    }
    block_sync();
}

}  // namespace mcrg
