// Body of the LDS-resident schedule: one workgroup advances the G lattices of
// one slot by up to n_chunk sweeps, measuring the sweeps that fall in each
// lattice's measurement phase.  Phases are loops over work items separated by
// workgroup barriers; every item is owned by exactly one thread, so no phase
// depends on which thread runs which item.
//
//   sweep      colour 0 then colour 1: one item per CB word, Metropolis in place
//   measure    level-0 operators + blocking to level 1 (CB -> RM), then levels
//              1..nl (RM), then one item per record entry adds the sample's
//              sums and products to the LDS partial record
//   flush      partial records are added to the per-bin records in HBM when a
//              lattice's sample index enters a new bin and at the chunk end
#pragma once

#include "levels.hpp"
#include "resident.hpp"

#ifndef MCRG_SYNC
#define MCRG_SYNC() __syncthreads()
#define MCRG_SYNC_OR(x) __syncthreads_or(x)
#endif

namespace mcrg {

#ifndef MCRG_HOST_EMULATION
// Sum over aligned groups of `width` lanes (a power of two <= 64); every lane of
// the wave must execute it.
__device__ __forceinline__ uint32_t group_sum(uint32_t v, uint32_t width)
{
    for (uint32_t off = 1; off < width; off <<= 1) v += __shfl_xor(v, (int)off, 64);
    return v;
}
__device__ __forceinline__ bool group_leader(uint32_t tid, uint32_t width) { return (tid & (width - 1u)) == 0u; }
#endif

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

__device__ __forceinline__ int64_t level_S(const uint32_t *cg, uint32_t L, uint32_t l, uint32_t a)
{
    const int64_t n = (int64_t)(L >> l) * (int64_t)(L >> l);
    const int64_t d = (int64_t)cg[l * 4u + a];
    return a < 2u ? 2 * n - 2 * d : n - 2 * d;
}

// Increment of record entry k for one sample (layout of record_size()).
__device__ __forceinline__ int64_t record_term(const uint32_t *cg, uint32_t L, uint32_t nl, uint32_t k)
{
    const uint32_t n_s = 3u * (nl + 1u), n_self = 9u * (nl + 1u), n_cross = 9u * nl;
    if (k < n_s) return level_S(cg, L, k / 3u, k % 3u);
    k -= n_s;
    if (k < n_self) {
        const uint32_t l = k / 9u, r = k % 9u;
        return level_S(cg, L, l, r % 3u) * level_S(cg, L, l, r / 3u);
    }
    k -= n_self;
    if (k < n_cross) {
        const uint32_t l = k / 9u + 1u, r = k % 9u;
        return level_S(cg, L, l, r % 3u) * level_S(cg, L, l - 1u, r / 3u);
    }
    k -= n_cross;
    const int64_t n0 = (int64_t)L * (int64_t)L;
    const int64_t M = n0 - 2 * (int64_t)cg[3];
    if (k == 0u) return M;
    if (k == 1u) return M < 0 ? -M : M;
    if (k == 2u) return M * M;
    return 1;
}

__device__ __forceinline__ bool in_run(const uint32_t *m, uint32_t it)
{
    return it < m[kMetaTherm] + m[kMetaMeas];
}
__device__ __forceinline__ bool in_meas(const uint32_t *m, uint32_t it)
{
    return it >= m[kMetaTherm] && it < m[kMetaTherm] + m[kMetaMeas];
}

__device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

// Adds `v` of item i (items of one lattice are contiguous, `per` per lattice)
// into counter slot `dst[g]`.  Groups of lanes that belong to the same lattice
// reduce first so that one lane per group touches LDS.
__device__ __forceinline__ void add_count(uint32_t *cnt, uint32_t stride, uint32_t i, uint32_t per,
                                          uint32_t n_items, uint32_t v, uint32_t tid)
{
    const uint32_t width = per < 64u ? per : 64u;
    v = group_sum(v, width);
    if (group_leader(tid, width) && i < n_items) atomicAdd(cnt + (i / per) * stride, v);
}

__device__ void run_slot(const Slot sl, const int32_t *__restrict__ slot_lats,
                         PackedLattice *__restrict__ lats, uint32_t n_chunk, uint32_t *lds,
                         uint32_t tid, uint32_t nt)
{
    const uint32_t L = (uint32_t)sl.L, G = (uint32_t)sl.G, nl = (uint32_t)sl.n_levels;
    const SlotLayout lay = slot_layout(L, G, nl);
    const uint32_t W = lay.W, rs = record_size(nl);
    const uint32_t cstride = (nl + 1u) * 4u;
    uint32_t *meta = lds + lay.meta;
    uint32_t *spins = lds + lay.spins;
    uint32_t *lvl = lds + lay.lvl;
    uint32_t *cnt = lds + lay.cnt;
    int64_t *rec = reinterpret_cast<int64_t *>(lds + lay.rec);

    for (uint32_t g = tid; g < G; g += nt) {
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
    MCRG_SYNC();
    for (uint32_t i = tid; i < G * 2u * W; i += nt) {
        const uint32_t g = i / (2u * W), j = i % (2u * W);
        const uint32_t li = meta[g * kMetaWords + kMetaIdx];
        spins[i] = li != 0xFFFFFFFFu ? lats[li].spins[j] : 0u;
    }
    for (uint32_t i = tid; i < G * rs; i += nt) rec[i] = 0;
    MCRG_SYNC();

    for (uint32_t it = 0; it < n_chunk; ++it) {
        int running = 0, measuring = 0;
        for (uint32_t g = tid; g < G; g += nt) {
            running |= in_run(meta + g * kMetaWords, it) ? 1 : 0;
            measuring |= in_meas(meta + g * kMetaWords, it) ? 1 : 0;
        }
        if (!MCRG_SYNC_OR(running)) break;

        // ---- sweep: colour 0, then colour 1
        for (uint32_t c = 0; c < 2u; ++c) {
            for (uint32_t i = tid; i < G * W; i += nt) {
                const uint32_t g = i / W, w = i % W;
                const uint32_t *m = meta + g * kMetaWords;
                if (!in_run(m, it)) continue;
                uint32_t *base = spins + g * 2u * W;
                uint32_t *own = base + c * W;
                const uint32_t *oth = base + (1u - c) * W;
                const Nbr4 n = cb_neighbours(oth, L, W, w, c);
                Accept acc;
                acc.thr0 = m[kMetaThr0];
                acc.thr1 = m[kMetaThr1];
                acc.always0 = m[kMetaAlw0];
                acc.always1 = m[kMetaAlw1];
                const uint32_t flip = metropolis_word(own[w], n.nu, n.nd, n.nl, n.nr, 0xFFFFFFFFu, acc,
                                                      m[kMetaK0], m[kMetaK1], w, m[kMetaT] + it, c);
                own[w] ^= flip;
            }
            MCRG_SYNC();
        }
        if (!MCRG_SYNC_OR(measuring)) continue;

        // ---- level 0: operators and blocking into level 1
        for (uint32_t i = tid; i < G * cstride; i += nt) cnt[i] = 0u;
        MCRG_SYNC();
        {
            const uint32_t per = 2u * W, n_items = G * per;
            const uint32_t rounds = ceil_div(n_items, nt);
            for (uint32_t r = 0; r < rounds; ++r) {
                const uint32_t i = tid + r * nt;
                uint32_t dnn = 0, dnnn = 0, dp = 0, nd = 0;
                if (i < n_items) {
                    const uint32_t g = i / per, j = i % per, c = j / W, w = j % W;
                    if (in_meas(meta + g * kMetaWords, it)) {
                        const uint32_t *base = spins + g * 2u * W;
                        cb_ops_item(base + c * W, base + (1u - c) * W, L, W, w, c, dnn, dnnn, dp, nd);
                    }
                }
                uint32_t *c0 = cnt;
                add_count(c0 + 0, cstride, i, per, n_items, dnn, tid);
                add_count(c0 + 1, cstride, i, per, n_items, dnnn, tid);
                add_count(c0 + 2, cstride, i, per, n_items, dp, tid);
                add_count(c0 + 3, cstride, i, per, n_items, nd, tid);
            }
            const uint32_t nb = nl > 0u ? cb_block_items(L) : 0u;  // no level-1 scratch without levels
            for (uint32_t i = tid; i < G * nb; i += nt) {
                const uint32_t g = i / nb, j = i % nb;
                const uint32_t *m = meta + g * kMetaWords;
                if (!in_meas(m, it)) continue;
                const uint32_t *base = spins + g * 2u * W;
                const uint32_t coin = tie_coin_word(j, m[kMetaT] + it, 0u, m[kMetaK0], m[kMetaK1]);
                lvl[g * lay.lvl_words + j] = cb_block_item(base, base + W, L, W, j, coin);
            }
        }
        MCRG_SYNC();

        // ---- levels 1..nl
        uint32_t off = 0;
        for (uint32_t l = 1; l <= nl; ++l) {
            const uint32_t Ll = L >> l;
            const uint32_t per = rm_ops_items(Ll), n_items = G * per;
            const uint32_t rounds = ceil_div(n_items, nt);
            for (uint32_t r = 0; r < rounds; ++r) {
                const uint32_t i = tid + r * nt;
                uint32_t dnn = 0, dnnn = 0, dp = 0;
                if (i < n_items) {
                    const uint32_t g = i / per, j = i % per;
                    if (in_meas(meta + g * kMetaWords, it))
                        rm_ops_item(lvl + g * lay.lvl_words + off, Ll, j, dnn, dnnn, dp);
                }
                uint32_t *cl = cnt + l * 4u;
                add_count(cl + 0, cstride, i, per, n_items, dnn, tid);
                add_count(cl + 1, cstride, i, per, n_items, dnnn, tid);
                add_count(cl + 2, cstride, i, per, n_items, dp, tid);
            }
            if (l < nl) {
                const uint32_t nb = rm_block_items(Ll);
                const uint32_t nxt = off + level_words(L, l);
                for (uint32_t i = tid; i < G * nb; i += nt) {
                    const uint32_t g = i / nb, j = i % nb;
                    const uint32_t *m = meta + g * kMetaWords;
                    if (!in_meas(m, it)) continue;
                    const uint32_t coin = tie_coin_word(j, m[kMetaT] + it, l, m[kMetaK0], m[kMetaK1]);
                    uint32_t *lg = lvl + g * lay.lvl_words;
                    lg[nxt + j] = rm_block_item(lg + off, Ll, j, coin);
                }
                off = nxt;
            }
            MCRG_SYNC();
        }

        // ---- the sample's sums and products into the partial records
        for (uint32_t i = tid; i < G * rs; i += nt) {
            const uint32_t g = i / rs, k = i % rs;
            const uint32_t *m = meta + g * kMetaWords;
            if (!in_meas(m, it)) continue;
            const uint32_t q = m[kMetaNMeas] + (it - m[kMetaTherm]);
            const uint32_t spb = m[kMetaSpb];
            if (q % spb == 0u && q > m[kMetaNMeas]) {
                lats[m[kMetaIdx]].records[(size_t)((q - 1u) / spb) * rs + k] += rec[i];
                rec[i] = 0;
            }
            rec[i] += record_term(cnt + g * cstride, L, nl, k);
        }
        MCRG_SYNC();
    }

    // ---- write back: partial records, spins, counters
    for (uint32_t i = tid; i < G * rs; i += nt) {
        const uint32_t g = i / rs, k = i % rs;
        const uint32_t *m = meta + g * kMetaWords;
        if (m[kMetaIdx] == 0xFFFFFFFFu) continue;
        const uint32_t total = m[kMetaTherm] + m[kMetaMeas];
        const uint32_t swept = total < n_chunk ? total : n_chunk;
        const uint32_t measured = swept > m[kMetaTherm] ? swept - m[kMetaTherm] : 0u;
        if (measured == 0u) continue;
        const uint32_t q_last = m[kMetaNMeas] + measured - 1u;
        lats[m[kMetaIdx]].records[(size_t)(q_last / m[kMetaSpb]) * rs + k] += rec[i];
    }
    for (uint32_t i = tid; i < G * 2u * W; i += nt) {
        const uint32_t g = i / (2u * W), j = i % (2u * W);
        const uint32_t li = meta[g * kMetaWords + kMetaIdx];
        if (li != 0xFFFFFFFFu) lats[li].spins[j] = spins[i];
    }
    for (uint32_t g = tid; g < G; g += nt) {
        const uint32_t *m = meta + g * kMetaWords;
        if (m[kMetaIdx] == 0xFFFFFFFFu) continue;
        PackedLattice &P = lats[m[kMetaIdx]];
        const uint32_t total = m[kMetaTherm] + m[kMetaMeas];
        const uint32_t swept = total < n_chunk ? total : n_chunk;
        const uint32_t measured = swept > m[kMetaTherm] ? swept - m[kMetaTherm] : 0u;
        P.t = m[kMetaT] + swept;
        P.n_measured = m[kMetaNMeas] + measured;
        P.therm_left = m[kMetaTherm] - (swept - measured);
        P.meas_left = m[kMetaMeas] - measured;
    }
    MCRG_SYNC();
}

}  // namespace mcrg
