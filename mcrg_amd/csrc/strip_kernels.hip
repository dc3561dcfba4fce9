// Strip schedule kernels (see strip.hpp for the structure).
//
//   k_strip_reset   clears a launch's workspace
//   k_strip         cooperative launch: workgroups 0 .. n_wg-1 each own a strip
//                   of R rows held bit-packed in LDS and run every sweep of the
//                   launch on it (checkerboard Metropolis replacing the Wolff
//                   step, src/ising.cpp:87-149), with the level-0 operators and
//                   the in-strip blockings of each measured sweep
//                   (src/lattice.cpp:102-120, src/mcrg.cpp:314-348); with
//                   split, workgroups n_wg .. 2 n_wg-1 measure the strips'
//                   samples instead; the last n_tail workgroups finish each
//                   sample's operators across strips and add its record terms
//                   (src/mcrg.cpp:72-98)
//
// Every random number comes from the same Philox counters as the other
// schedules (device_common.hpp), so a chain's spins and records are
// bit-identical to the HBM schedule's and the oracle's.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "levels.hpp"
#include "strip.hpp"

namespace mcrg {

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One (data, seq) word: a single 64-bit store with agent scope, so a reader
// on any compute unit sees either the old word or the new one whole.
__device__ __forceinline__ void ll_put(uint64_t *p, uint32_t v, uint32_t seq)
{
    __hip_atomic_store(p, (uint64_t)v | ((uint64_t)seq << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t ll_peek(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Polls *p until its sequence is `seq` and returns the word.  Gives up (the
// returned word's sequence is then not `seq`) once the launch's error word is
// set, by this thread after kStripWaitTicks of the 100 MHz clock or by any
// other, so that every workgroup reaches the end of the kernel.  Not inlined:
// the slow path stays out of the sweep's registers.
__device__ __noinline__ uint64_t ll_wait(const uint64_t *p, uint32_t seq, uint32_t *err, uint32_t sleep)
{
    const uint64_t t0 = (uint64_t)wall_clock64();
    for (uint32_t k = 1;; ++k) {
        const uint64_t v = ll_peek(p);
        if ((uint32_t)(v >> 32) == seq) return v;
        if ((k & 7u) == 0u) {
            if (ld_relaxed(err) != 0u) return 0;
            if ((uint64_t)wall_clock64() - t0 > kStripWaitTicks) {
                __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return 0;
            }
        }
        for (uint32_t z = 0; z < sleep; ++z) __builtin_amdgcn_s_sleep(1);
    }
}

// The data of (data, seq) word *p once it carries `seq`, polling every
// `sleep` x 64 clocks; `bad` is set when the wait gave up.
__device__ __forceinline__ uint32_t ll_get(const uint64_t *p, uint32_t seq, uint32_t *err, bool &bad, uint32_t sleep)
{
    uint64_t v = ll_peek(p);
    if ((uint32_t)(v >> 32) != seq) {
        v = ll_wait(p, seq, err, sleep);
        if ((uint32_t)(v >> 32) != seq) bad = true;
    }
    return (uint32_t)v;
}

// Thread 0 polls the 32-bit *flag until it is >= want.  Relaxed loads: what
// depends on the poll is this workgroup's later stores into memory the flag's
// writer has finished reading, not data it wrote.  False after an error.
__device__ __noinline__ bool flag_wait(const uint32_t *flag, uint32_t want, uint32_t *err)
{
    const uint64_t t0 = (uint64_t)wall_clock64();
    while (ld_relaxed(flag) < want) {
        if (ld_relaxed(err) != 0u) return false;
        __builtin_amdgcn_s_sleep(1);
        if ((uint64_t)wall_clock64() - t0 > kStripWaitTicks) {
            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    return true;
}

// Adds the wave's sums of n <= 4 per-lane counts into dst[0..n-1] of LDS (one
// atomic per count and wave, zeros skipped).  No barrier: the caller orders
// the reads of dst after every wave's add with one of its own.
// Sum over the wave, valid in lane 0: DPP row operations within rows of 16
// lanes, a swizzle across the two halves of each 32, a readlane across 32.
__device__ __forceinline__ uint32_t wave_sum0(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    v += (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);                // xor 16
    return v + (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
}

__device__ __forceinline__ void wave_add(uint32_t *dst, uint32_t n, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3)
{
    if (__ballot((c0 | c1 | c2 | c3) != 0u) == 0ull) return;  // the wave has nothing to add
    const uint32_t v[4] = {wave_sum0(c0), wave_sum0(c1), n > 2u ? wave_sum0(c2) : 0u, n > 3u ? wave_sum0(c3) : 0u};
    if ((threadIdx.x & 63u) == 0u)
        for (uint32_t k = 0; k < n; ++k)
            if (v[k]) atomicAdd(dst + k, v[k]);
}

__device__ __forceinline__ void load64b(const int8_t *p, uint32_t d[16])
{
    const uint4 *v = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint4 x = v[i];
        d[4 * i + 0] = x.x;
        d[4 * i + 1] = x.y;
        d[4 * i + 2] = x.z;
        d[4 * i + 3] = x.w;
    }
}

// The 64 int8 spins of a 64-column chunk from the spin bits of its even
// (e) and odd (o) positions: byte 2j takes bit j of e, byte 2j+1 bit j of o
// (bit 1 = spin -1 = 0xFF, bit 0 = +1 = 0x01).
__device__ __forceinline__ void store64b(int8_t *p, uint32_t e, uint32_t o)
{
    uint4 *v = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint32_t d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t j = 4u * (uint32_t)i + (uint32_t)k;  // dword j holds bytes 4j .. 4j+3
            const uint32_t b0 = (e >> (2u * j)) & 1u, b1 = (o >> (2u * j)) & 1u;
            const uint32_t b2 = (e >> (2u * j + 1u)) & 1u, b3 = (o >> (2u * j + 1u)) & 1u;
            d[k] = (b0 ? 0xFFu : 0x01u) | ((b1 ? 0xFFu : 0x01u) << 8) | ((b2 ? 0xFFu : 0x01u) << 16) |
                   ((b3 ? 0xFFu : 0x01u) << 24);
        }
        v[i] = make_uint4(d[0], d[1], d[2], d[3]);
    }
}

__global__ void k_strip_reset(uint32_t *ws, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ws[i] = 0u;
}

// Level-0 operator counts of one checkerboard word: colour c, local row r
// (global row y), word q.  own / oth hold rows y0-1 .. y0+R as local rows 0 .. R+1.
__device__ __forceinline__ void strip_ops0(const uint32_t *own, const uint32_t *oth, uint32_t wpr, uint32_t r,
                                           uint32_t q, uint32_t y, uint32_t c, uint32_t &dnn, uint32_t &dnnn,
                                           uint32_t &dp, uint32_t &ndown)
{
    const uint32_t m = wpr - 1u, row = r * wpr;
    const uint32_t s = own[row + q];
    const uint32_t nu = oth[row - wpr + q], nd = oth[row + wpr + q], o = oth[row + q];
    const uint32_t nx = own[row + wpr + q];
    // the words beside o and nx that hold the wrap-around sites: q - 1 on
    // p = 0 rows, q + 1 on p = 1 rows (one load each, before any wait)
    const uint32_t p = (y + c) & 1u, qs = (q + 2u * p - 1u) & m;
    const uint32_t os = oth[row + qs], nxs = own[row + wpr + qs];
    const uint32_t nr = p == 0u ? o : (o >> 1) | (os << 31);
    const uint32_t nl = p == 0u ? (o << 1) | (os >> 31) : o;
    const uint32_t dr = p == 0u ? nx : (nx >> 1) | (nxs << 31);
    const uint32_t dl = p == 0u ? (nx << 1) | (nxs >> 31) : nx;
    // every nearest-neighbour bond has one colour-0 end
    if (c == 0u) dnn += (uint32_t)(popc(s ^ nu) + popc(s ^ nd) + popc(s ^ nl) + popc(s ^ nr));
    dnnn += (uint32_t)(popc(s ^ dr) + popc(s ^ dl));
    dp += (uint32_t)popc(xor3(s, nr, nd) ^ dr);
    ndown += (uint32_t)popc(s);
}

// Operator counts of a row-major row pair (row a above row d, rows of >= 32
// sites) at one word, from a[q], a[q+1], d[q], d[q+1] (wrapping within the
// row): the row-major item of levels.hpp.
__device__ __forceinline__ void pair_ops(uint32_t a, uint32_t an, uint32_t d, uint32_t dn, uint32_t &dnn,
                                         uint32_t &dnnn, uint32_t &dp)
{
    const uint32_t r = (a >> 1) | (an << 31), g = (d >> 1) | (dn << 31);
    dnn += (uint32_t)(popc(a ^ r) + popc(a ^ d));
    dnnn += (uint32_t)(popc(a ^ g) + popc(r ^ d));
    dp += (uint32_t)popc(xor3(a, r, d) ^ g);
}

// Barrier for LDS only: unlike __syncthreads it does not wait for this
// wave's stores to device memory (agent-scope stores that go all the way to
// memory) to complete.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup-wide OR of `bad` through one LDS word: a thread whose wait gave
// up stores `tag` (never 0, different for every call of a launch), and after
// an LDS-only barrier every thread reads the word.  Unlike __syncthreads_or
// it neither reads the dispatch packet from memory nor waits for the wave's
// agent-scope stores to complete.  A later call's store cannot overtake an
// earlier call's read: no thread gets past a call whose word says failure.
__device__ __forceinline__ bool strip_any(bool bad, uint32_t *word, uint32_t tag)
{
    if (bad) *word = tag;
    lds_barrier();
    return *word == tag;
}

__device__ __forceinline__ uint32_t lg2(uint32_t v) { return (uint32_t)__builtin_ctz(v); }

// LDS of a strip workgroup, in 32-bit words: both colours of rows y0-1 ..
// y0+R (colour c at cb0 + c * cbs), the in-strip levels, their tie coins
// (same offsets), the pending sample's first rows (frow), last rows (lrow),
// the lower strip's first rows (dfirst) and counts (pend), scratch, and with
// overlap the copy of both colours the pending sample is measured from.
struct StripLds {
    uint32_t *cb0, *lvl, *coin, *frow, *lrow, *dfirst, *pend, *red, *snap;
    uint32_t cbs;
    uint32_t w;  // the strip
};

__device__ __forceinline__ StripLds strip_lds(const StripShape &sh, uint32_t *lds, uint32_t w)
{
    StripLds m;
    m.w = w;
    m.cbs = (sh.R + 2u + 2u * sh.halo) * sh.wpr;
    m.cb0 = lds;
    m.lvl = lds + 2u * m.cbs;
    uint32_t lv = 0;
    for (uint32_t l = 1; l <= sh.lam; ++l) lv += (sh.R >> l) * strip_row_words(sh.L, l);
    m.coin = m.lvl + lv;
    m.frow = m.coin + lv;
    m.lrow = m.frow + sh.first_words;
    m.dfirst = m.lrow + sh.first_words;
    m.pend = m.dfirst + sh.first_words;
    m.red = m.pend + kStripCntWords;
    m.snap = m.red + kRedWords;
    return m;
}

// The tie coins of every in-strip blocking of the sample at sweep t, into
// m.coin: level l's words are the strip's range of that level's row-major
// words, (y0 >> l) * wl .. + (R >> l) * wl, and one Philox call gives four
// consecutive coin words (the stream of tie_coin_word).  The calls of all
// levels are spread over the workgroup at once.
template <int NT>
__device__ void strip_coins(const StripArgs &a, const StripLds &m, uint32_t t, uint32_t k0, uint32_t k1)
{
    const StripShape &sh = a.sh;
    const uint32_t L = sh.L, R = sh.R, y0 = m.w * R;
    uint32_t calls = 0;
    for (uint32_t l = 1; l <= sh.lam; ++l) {
        const uint32_t n = (R >> l) * strip_row_words(L, l), base = (y0 >> l) * strip_row_words(L, l);
        calls += ((base + n - 1u) >> 2) - (base >> 2) + 1u;
    }
    // on the workgroup's last threads, so that the first ones' Philox calls
    // for the next update run beside them
    for (uint32_t j = NT - 1u - threadIdx.x; j < calls; j += NT) {
        uint32_t l = 1, g = j, off = 0, n = 0, base = 0;
        for (;; ++l) {
            n = (R >> l) * strip_row_words(L, l);
            base = (y0 >> l) * strip_row_words(L, l);
            const uint32_t c = ((base + n - 1u) >> 2) - (base >> 2) + 1u;
            if (g < c) break;
            g -= c;
            off += n;
        }
        const uint32_t grp = (base >> 2) + g;
        const u32x4 r = philox(grp, t, 0u, (kTagTie << 28) | ((l - 1u) << 16), k0, k1);
        const uint32_t w4[4] = {r.v0, r.v1, r.v2, r.v3};
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            const uint32_t gi = 4u * grp + k;
            if (gi >= base && gi < base + n) m.coin[off + (gi - base)] = w4[k];
        }
    }
}

// Synchronisation of the threads that run strip_levels: the workgroup's LDS
// barrier, or for one wave (whose LDS accesses complete in order) a wave
// barrier that keeps the compiler from moving LDS accesses across it.
template <bool WAVE>
__device__ __forceinline__ void group_sync()
{
    if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        lds_barrier();
    }
}

// The in-strip blocked levels 1..lam of one sample: blockings (coins from
// m.coin), the counts of every row pair inside the strip into pend, and each
// level's first and last row kept for the boundary bonds.  Run by N threads,
// tid = 0 .. N-1: the workgroup, or (WAVE) one wave.
template <uint32_t N, bool WAVE>
__device__ void strip_levels(const StripArgs &a, const StripLds &m, const uint32_t *src, uint32_t tid)
{
    const StripShape &sh = a.sh;
    const uint32_t L = sh.L, R = sh.R, wpr = sh.wpr, sw = lg2(wpr);
    const uint32_t *cb0 = src, *cb1 = src + m.cbs;
    // level 1: block (bx, by) holds bit bx of the colour-0 and colour-1 words
    // of rows 2by and 2by+1
    uint32_t *cur = m.lvl;
    const uint32_t *coin = m.coin;
    for (uint32_t j = tid; j < (R >> 1) * wpr; j += N) {
        const uint32_t lb = j >> sw, qq = j & (wpr - 1u);
        const uint32_t r0 = ((1u + 2u * lb) << sw) + qq, r1 = r0 + wpr;
        const uint32_t b1 = cb0[r0], b2 = cb1[r0], b3 = cb0[r1], b4 = cb1[r1];
        const uint32_t maj3 = (b1 & b2 & (b3 | b4)) | (b3 & b4 & (b1 | b2));
        const uint32_t tie = ~(xor3(b1, b2, b3) ^ b4) & ~(b1 & b2 & b3 & b4) & (b1 | b2 | b3 | b4);
        cur[j] = maj3 | (tie & coin[j]);
    }
    group_sync<WAVE>();
    for (uint32_t l = 1; l <= sh.lam; ++l) {
        const uint32_t Rl = R >> l, wl = strip_row_words(L, l), sl = lg2(wl);
        uint32_t d0 = 0, d1 = 0, d2 = 0;
        for (uint32_t j = tid; j < (Rl - 1u) << sl; j += N) {
            const uint32_t q = j & (wl - 1u), qn = (q + 1u) & (wl - 1u);
            const uint32_t *ar = cur + (j - q), *dr = ar + wl;
            pair_ops(ar[q], ar[qn], dr[q], dr[qn], d0, d1, d2);
        }
        // the first and last rows and the next blocking on other threads than
        // the pair counts, so that their latencies overlap
        for (uint32_t j = (tid + N / 2u) % N; j < wl; j += N) {
            m.frow[sh.row_off[l] + j] = cur[j];
            m.lrow[sh.row_off[l] + j] = cur[((Rl - 1u) << sl) + j];
        }
        wave_add(m.pend + 4u * l, 3u, d0, d1, d2, 0u);
        if (l < sh.lam) {
            uint32_t *nxt = cur + (Rl << sl);
            coin += Rl << sl;  // level l + 1's coins follow level l's
            const uint32_t wo = wl >> 1, so = sl - 1u;
            for (uint32_t j = N - 1u - tid; j < (Rl >> 1) << so; j += N) {
                const uint32_t jo = j >> so, qo = j & (wo - 1u);
                const uint32_t c = coin[j];
                uint32_t o = 0;
#pragma unroll
                for (uint32_t h = 0; h < 2u; ++h) {
                    const uint32_t w0 = cur[((2u * jo) << sl) + 2u * qo + h], w1 = cur[((2u * jo + 1u) << sl) + 2u * qo + h];
                    o |= block16(w0, w1, (c >> (16u * h)) & 0xFFFFu) << (16u * h);
                }
                nxt[j] = o;
            }
            group_sync<WAVE>();
            cur = nxt;
        }
    }
    group_sync<WAVE>();
}

// Measurement of a sample in a strip workgroup from its spins at src (both
// colours of rows y0-1 .. y0+R, colour c at src + c * cbs): level-0 operators
// of the strip's rows, then the in-strip levels (whose coins strip_coins drew
// during the sample's colour-1 exchange).  The sample stays pending in LDS
// until its boundary bonds are counted (strip_publish).  Only LDS barriers:
// neighbour words loaded before it stay in flight.
template <int NT>
__device__ void strip_measure(const StripArgs &a, const StripLds &m, const uint32_t *src)
{
    const StripShape &sh = a.sh;
    const uint32_t R = sh.R, wpr = sh.wpr, y0 = m.w * R, sw = lg2(wpr), sc = lg2(R * wpr);
    // pend is zero: cleared at the start of the launch and by the previous
    // sample's publication, which ended with a barrier
    uint32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    // Small strips (level-1 blocking of at most 4 words per lane): the whole
    // chain of blocked levels on wave 0, with wave barriers only, while the
    // other waves count the level-0 operators.
    const bool one_wave = NT >= 128 && sh.lam > 0u && (R >> 1) * wpr <= 256u;
    if (one_wave && threadIdx.x < 64u) {
        strip_levels<64u, true>(a, m, src, threadIdx.x);
    } else {
        const uint32_t t0 = one_wave ? 64u : 0u;
        for (uint32_t j = threadIdx.x - t0; j < 2u * R * wpr; j += NT - t0) {
            const uint32_t c = j >> sc, jj = j & ((1u << sc) - 1u);
            const uint32_t r = 1u + (jj >> sw), qq = jj & (wpr - 1u);
            strip_ops0(src + c * m.cbs, src + (c ^ 1u) * m.cbs, wpr, r, qq, y0 + r - 1u, c, d0, d1, d2, d3);
        }
        wave_add(m.pend, 4u, d0, d1, d2, d3);
    }
    if (!one_wave && sh.lam > 0u) strip_levels<NT, false>(a, m, src, threadIdx.x);
    else lds_barrier();
}

// Sends the pending sample's first rows to the strip above (the words the
// caller's exchange reads back from the strip below into dfirst).
__device__ __forceinline__ void strip_send_first(const StripArgs &a, const StripLds &m, uint32_t seq, uint32_t j)
{
    uint64_t *xf = reinterpret_cast<uint64_t *>(a.ws + a.sh.o_xfirst) + (size_t)blockIdx.x * a.sh.first_words;
    ll_put(xf + j, m.frow[j], seq);
}

__device__ __forceinline__ uint32_t strip_recv_first(const StripArgs &a, uint32_t seq, uint32_t j, uint32_t *err,
                                                     bool &bad)
{
    const uint32_t dn = (blockIdx.x + 1u) & (a.sh.n_wg - 1u);
    const uint64_t *xf = reinterpret_cast<const uint64_t *>(a.ws + a.sh.o_xfirst) + (size_t)dn * a.sh.first_words;
    return ll_get(xf + j, seq, err, bad, a.strip_sleep);
}

// Completes pending sample q (the lower strip's first rows are in dfirst):
// the bonds between this strip's last blocked rows and them, then the counts
// and the level-lam row to the sample's ring slot.  False on an error.
template <int NT>
__device__ bool strip_publish(const StripArgs &a, const StripLds &m, uint32_t q)
{
    const StripShape &sh = a.sh;
    uint32_t *err = a.ws + sh.o_err;
    // one word of one level per thread: level l's row is words
    // row_off[l] .. + wl of the first-rows block (levels 1..lam, halving)
    for (uint32_t j = NT - 1u - threadIdx.x; j < sh.first_words; j += NT) {
        uint32_t l = 1u, o = 0u, wl = strip_row_words(sh.L, 1u);
        while (l < sh.lam && j >= o + wl) {
            o += wl;
            wl >>= 1;
            ++l;
        }
        const uint32_t i = j - o, in = (i + 1u) & (wl - 1u);
        uint32_t d0 = 0, d1 = 0, d2 = 0;
        pair_ops(m.lrow[j], m.lrow[o + in], m.dfirst[j], m.dfirst[o + in], d0, d1, d2);
        uint32_t *p = m.pend + 4u * l;
        if (d0) atomicAdd(p, d0);
        if (d1) atomicAdd(p + 1, d1);
        if (d2) atomicAdd(p + 2, d2);
    }
    const uint32_t s = q % kStripRing, gen = q / kStripRing, seq = q + 1u;
    // The tail must have finished with this slot's previous sample.  The
    // slots' consumed counts seen last are cached in LDS (red[0 .. ring)); when
    // the cache says no, thread 0 reads all of them at once and polls only if
    // this slot is still behind, so most samples pay no round trip to memory.
    if (gen > 0u && threadIdx.x == 0) {
        uint32_t *seen = m.red;
        bool ok = true;
        if (seen[s] < gen) {
            uint32_t v[kStripRing];
#pragma unroll
            for (uint32_t i = 0; i < kStripRing; ++i) v[i] = ld_relaxed(a.ws + sh.o_consumed + i * kStripFlagStride);
#pragma unroll
            for (uint32_t i = 0; i < kStripRing; ++i) seen[i] = v[i];
            if (seen[s] < gen) {
                ok = flag_wait(a.ws + sh.o_consumed + s * kStripFlagStride, gen, err);
                seen[s] = gen;
            }
        }
        m.red[kRedWait] = ok ? 1u : 0u;
    }
    lds_barrier();  // every wave's counts are in pend, and the wait's answer
    if (gen > 0u && m.red[kRedWait] == 0u) return false;
    uint64_t *out = reinterpret_cast<uint64_t *>(a.ws + sh.o_ring) + (size_t)(s * sh.n_wg + m.w) * sh.per_wg;
    for (uint32_t j = threadIdx.x; j < sh.per_wg; j += NT) {
        if (j < sh.cnt_words) {
            ll_put(out + j, m.pend[j], seq);
            m.pend[j] = 0u;  // for the next sample (each word is read by this thread only)
        } else {
            ll_put(out + j, m.frow[sh.row_off[sh.lam] + (j - sh.cnt_words)], seq);
        }
    }
    lds_barrier();  // pend and frow are rewritten by the next measurement
    return true;
}

template <int NT>
__device__ void strip_body(const StripArgs &a, uint32_t *lds)
{
    const StripShape &sh = a.sh;
    const HbmLattice &lat = *a.lat;
    const uint32_t L = sh.L, R = sh.R, wpr = sh.wpr, nwg = sh.n_wg, sw = lg2(wpr);
    const uint32_t w = blockIdx.x, y0 = w * R;
    const uint32_t up = (w + nwg - 1u) & (nwg - 1u), dn = (w + 1u) & (nwg - 1u);
    const StripLds m = strip_lds(sh, lds, w);
    uint32_t *const cb0 = m.cb0;
    const uint32_t cbs = m.cbs, fw = sh.first_words;
    uint32_t *ws = a.ws;
    uint32_t *err = ws + sh.o_err;
    uint64_t *xhalo = reinterpret_cast<uint64_t *>(ws + sh.o_xhalo);
    // rows y0-1 .. y0+R from the int8 lattice; colour c of row y sits at the
    // chunk positions of parity (y + c) & 1
    for (uint32_t j = threadIdx.x; j < (R + 2u) * wpr; j += NT) {
        const uint32_t r = j >> sw, q = j & (wpr - 1u);
        const uint32_t y = (y0 + r + L - 1u) & (L - 1u);
        uint32_t d[16];
        load64b(lat.spins + (size_t)y * L + 64u * q, d);
        const uint32_t ce = y & 1u;  // the colour at even positions
        cb0[ce * cbs + j] = gather_parity(d, 0u);
        cb0[(ce ^ 1u) * cbs + j] = gather_parity(d, 1u);
    }
    for (uint32_t i = threadIdx.x; i < kStripRing; i += NT) m.red[i] = 0u;  // consumed counts seen
    for (uint32_t i = threadIdx.x; i < kStripCntWords; i += NT) m.pend[i] = 0u;
    if (threadIdx.x == 0) m.red[kRedAny] = 0u;
    __syncthreads();
    const Accept acc = lat.acc;
    const uint32_t k0 = lat.k0, k1 = lat.k1, mk = wpr - 1u;
    uint32_t vote = 0;
    // With at most one word per thread (R wpr <= NT), thread j's word of
    // either colour is word y0 wpr + j, and the first two Philox calls of its
    // next update are drawn while the previous exchange is in flight.
    const bool one = R * wpr <= (uint32_t)NT;
    const uint32_t jw = threadIdx.x, my_word = y0 * wpr + jw;
    u32x4 pr{0u, 0u, 0u, 0u}, pref{0u, 0u, 0u, 0u};
    if (one && jw < R * wpr && a.n_sweeps > 0u) metropolis_rng(my_word, a.t0, 0u, k0, k1, pr, pref);
    bool ok = true;
    for (uint32_t it = 0; it < a.n_sweeps && ok; ++it) {
        const uint32_t t = a.t0 + it;
        for (uint32_t c = 0; c < 2u && ok; ++c) {
            uint32_t *own = cb0 + c * cbs;
            const uint32_t *oth = cb0 + (c ^ 1u) * cbs;
            for (uint32_t j = threadIdx.x; j < R * wpr; j += NT) {
                const uint32_t r = 1u + (j >> sw), q = j & mk, y = y0 + r - 1u, row = r << sw;
                // o and the word beside it that holds the wrap-around site
                // (q - 1 on p = 0 rows, q + 1 on p = 1 rows): one load each
                const uint32_t p = (y + c) & 1u;
                const uint32_t o = oth[row + q], os = oth[row + ((q + 2u * p - 1u) & mk)];
                const uint32_t nr = p == 0u ? o : (o >> 1) | (os << 31);
                const uint32_t nl = p == 0u ? (o << 1) | (os >> 31) : o;
                if (one)
                    own[row + q] ^= metropolis_word_drawn(own[row + q], oth[row - wpr + q], oth[row + wpr + q], nl, nr,
                                                          acc, k0, k1, y * wpr + q, t, c, pr, pref);
                else
                    own[row + q] ^= metropolis_word(own[row + q], oth[row - wpr + q], oth[row + wpr + q], nl, nr,
                                                    0xFFFFFFFFu, acc, k0, k1, y * wpr + q, t, c);
            }
            lds_barrier();
            // this colour's first and last rows to the neighbours, theirs back:
            // local row 0 = the last row of the strip above, R + 1 = the first
            // of the strip below.  The colour-0 exchange also carries the
            // previous sample's first blocked rows upwards; while the colour-1
            // exchange is in flight the sample's tie coins are drawn.
            const uint32_t seq = 2u * it + c + 1u;
            // the pending sample's first rows travel with the colour-0
            // exchange, or with overlap (the sample is measured during the
            // colour-0 exchange) with the colour-1 exchange
            const bool carry = c == (sh.overlap ? 1u : 0u) && sh.measure && it > 0u && fw > 0u;
            uint64_t *mine = xhalo + (size_t)(c * nwg + w) * 2u * wpr;
            for (uint32_t j = threadIdx.x; j < 2u * wpr; j += NT)
                ll_put(mine + j, own[(j < wpr ? 1u : R) * wpr + (j & mk)], seq);
            if (carry)
                for (uint32_t j = threadIdx.x; j < fw; j += NT) strip_send_first(a, m, it, j);
            // the neighbours' words (2 wpr <= NT and fw <= NT: one of each per
            // thread at most) are loaded first and checked last
            const uint32_t j = threadIdx.x;
            const uint64_t *src_h = j < wpr ? xhalo + (size_t)(c * nwg + up) * 2u * wpr + wpr + j
                                            : xhalo + (size_t)(c * nwg + dn) * 2u * wpr + (j - wpr);
            const uint64_t *src_f = reinterpret_cast<const uint64_t *>(ws + sh.o_xfirst) + (size_t)dn * fw + j;
            const bool get_h = j < 2u * wpr, get_f = carry && j < fw;
            uint64_t vh = get_h ? ll_peek(src_h) : 0ull, vf = get_f ? ll_peek(src_f) : 0ull;
            // meanwhile: with overlap, the previous sample's measurement (in
            // the colour-0 exchange); the next update's first Philox calls;
            // and in the colour-1 exchange of a measured sweep the sample's
            // tie coins
            if (sh.overlap && c == 0u && it > 0u) strip_measure<NT>(a, m, m.snap);
            if (one && jw < R * wpr && (c == 0u || it + 1u < a.n_sweeps))
                metropolis_rng(my_word, c == 0u ? t : t + 1u, c ^ 1u, k0, k1, pr, pref);
            if (c == 1u && sh.measure && sh.lam > 0u) strip_coins<NT>(a, m, t, k0, k1);
            bool bad = false;
            if (get_h) {
                if ((uint32_t)(vh >> 32) != seq) {
                    vh = ll_wait(src_h, seq, err, a.strip_sleep);
                    if ((uint32_t)(vh >> 32) != seq) bad = true;
                }
                own[j < wpr ? j : (R + 1u) * wpr + (j - wpr)] = (uint32_t)vh;
            }
            if (get_f) {
                if ((uint32_t)(vf >> 32) != it) {
                    vf = ll_wait(src_f, it, err, a.strip_sleep);
                    if ((uint32_t)(vf >> 32) != it) bad = true;
                }
                m.dfirst[j] = (uint32_t)vf;
            }
            ok = !strip_any(bad, m.red + kRedAny, ++vote);
            if (ok && c == (sh.overlap ? 1u : 0u) && sh.measure && it > 0u) ok = strip_publish<NT>(a, m, it - 1u);
        }
        if (ok && sh.measure) {
            if (sh.overlap) {
                // the spins of this sample, measured during the next exchange
                for (uint32_t i = threadIdx.x; i < 2u * cbs; i += NT) m.snap[i] = cb0[i];
                lds_barrier();
            } else {
                strip_measure<NT>(a, m, cb0);
            }
        }
    }
    if (ok && sh.measure && a.n_sweeps > 0u) {
        if (sh.overlap) strip_measure<NT>(a, m, m.snap);
        // the last sample's boundary rows: one more exchange of first rows
        bool bad = false;
        if (fw > 0u) {
            for (uint32_t j = threadIdx.x; j < fw; j += NT) strip_send_first(a, m, a.n_sweeps, j);
            for (uint32_t j = threadIdx.x; j < fw; j += NT)
                m.dfirst[j] = strip_recv_first(a, a.n_sweeps, j, err, bad);
        }
        ok = !strip_any(bad, m.red + kRedAny, ++vote);
        if (ok) ok = strip_publish<NT>(a, m, a.n_sweeps - 1u);
    }
    if (!ok) return;  // the host reports the error word; the lattice is not written back
    for (uint32_t j = threadIdx.x; j < R * wpr; j += NT) {
        const uint32_t r = 1u + (j >> sw), q = j & mk, y = y0 + r - 1u;
        const uint32_t ce = y & 1u;
        store64b(lat.spins + (size_t)y * L + 64u * q, cb0[ce * cbs + r * wpr + q], cb0[(ce ^ 1u) * cbs + r * wpr + q]);
    }
    if (w == 0u && threadIdx.x == 0u) {
        lat.state[0] = a.t0 + a.n_sweeps;
        lat.state[1] = a.q0 + (sh.measure ? a.n_sweeps : 0u);
    }
}

// The flips of checkerboard word q of local row `row` (global row y) of
// colour c, own / oth its colour's and the other colour's rows, with the
// first two Philox calls drawn beforehand.
__device__ __forceinline__ uint32_t strip_flip(const uint32_t *own, const uint32_t *oth, uint32_t row, uint32_t q,
                                               uint32_t wpr, uint32_t y, uint32_t c, const Accept &acc, uint32_t k0,
                                               uint32_t k1, uint32_t t, const u32x4 &r, const u32x4 &ref)
{
    // o and the word beside it that holds the wrap-around site (q - 1 on
    // p = 0 rows, q + 1 on p = 1 rows): one load each, before any wait
    const uint32_t mk = wpr - 1u, p = (y + c) & 1u;
    const uint32_t o = oth[row + q], os = oth[row + ((q + 2u * p - 1u) & mk)];
    const uint32_t nr = p == 0u ? o : (o >> 1) | (os << 31);
    const uint32_t nl = p == 0u ? (o << 1) | (os >> 31) : o;
    return metropolis_word_drawn(own[row + q], oth[row - wpr + q], oth[row + wpr + q], nl, nr, acc, k0, k1,
                                 y * wpr + q, t, c, r, ref);
}

// The halo-row form of strip_body (sh.halo).  Rows y0-2 .. y0+R+1 are local
// rows 0 .. R+3; thread j updates colour-0 word j of rows 2 .. R+1 (the
// strip's own) and colour-1 words j and j + NT of rows 1 .. R+2.  One sweep t:
//   * colour 0 on rows 2 .. R+1;
//   * the strip's first two and last two colour-0 rows to the neighbours (one
//     exchange per sweep, buffer t & 1).  While they travel, when the strip
//     measures its own samples: sample t-2 is published (the lower strip's
//     first blocked rows of it were sent in the previous sweep), sample t-1 is
//     measured from its copy and its first blocked rows sent, the tie coins of
//     sample t are drawn; and the next two updates' first Philox calls;
//   * the neighbours' rows into local rows 0, 1, R+2, R+3;
//   * colour 1 on rows 1 .. R+2.  Rows 1 and R+2 belong to the neighbours,
//     whose decisions for them come from the same counters, so afterwards
//     rows 1 .. R+2 hold sample t whole and the next colour-0 update needs no
//     exchange;
//   * when measuring, the copy of sample t: in LDS after colour 1, or with
//     split to the strip's measuring workgroup (slot t mod kStripSnapSlots,
//     once that workgroup has copied the slot's previous sample), each word
//     as soon as it is final: colour 0 after the exchange, colour 1 from the
//     update's registers.
// A sample's first blocked rows are read one sweep after they are sent, from
// first_slots slots per strip: a strip may send the next two samples' before
// its upper neighbour has read them.
template <int NT>
__device__ void strip_body_halo(const StripArgs &a, uint32_t *lds)
{
    const StripShape &sh = a.sh;
    const HbmLattice &lat = *a.lat;
    const uint32_t L = sh.L, R = sh.R, wpr = sh.wpr, nwg = sh.n_wg, sw = lg2(wpr), mk = wpr - 1u;
    const uint32_t w = blockIdx.x, y0 = w * R;
    const uint32_t up = (w + nwg - 1u) & (nwg - 1u), dn = (w + 1u) & (nwg - 1u);
    const StripLds m = strip_lds(sh, lds, w);
    uint32_t *const cb0 = m.cb0, *const cb1 = m.cb0 + m.cbs;
    const uint32_t cbs = m.cbs, fw = sh.first_words, hw = 4u * wpr, fs = sh.first_slots;
    uint32_t *err = a.ws + sh.o_err;
    uint64_t *xhalo = reinterpret_cast<uint64_t *>(a.ws + sh.o_xhalo);
    uint64_t *xfirst = reinterpret_cast<uint64_t *>(a.ws + sh.o_xfirst);
    // rows y0-2 .. y0+R+1 from the int8 lattice
    for (uint32_t j = threadIdx.x; j < (R + 4u) * wpr; j += NT) {
        const uint32_t r = j >> sw, q = j & mk;
        const uint32_t y = (y0 + r + L - 2u) & (L - 1u);
        uint32_t d[16];
        load64b(lat.spins + (size_t)y * L + 64u * q, d);
        const uint32_t ce = y & 1u;  // the colour at even positions
        cb0[ce * cbs + j] = gather_parity(d, 0u);
        cb0[(ce ^ 1u) * cbs + j] = gather_parity(d, 1u);
    }
    for (uint32_t i = threadIdx.x; i < kStripRing; i += NT) m.red[i] = 0u;  // consumed counts seen
    for (uint32_t i = threadIdx.x; i < kStripCntWords; i += NT) m.pend[i] = 0u;
    if (threadIdx.x == 0) m.red[kRedAny] = 0u;
    __syncthreads();
    const Accept acc = lat.acc;
    const uint32_t k0 = lat.k0, k1 = lat.k1, n = a.n_sweeps;
    const bool own = sh.measure && !sh.split;  // this workgroup measures its samples
    const uint32_t jw = threadIdx.x, jr = jw >> sw, q = jw & mk;
    const uint32_t j2 = jw + NT, jr2 = j2 >> sw, q2 = j2 & mk;
    const bool has0 = jw < R * wpr, has1 = jw < (R + 2u) * wpr, has2 = j2 < (R + 2u) * wpr;
    // thread jw's colour-0 word: local row 2 + jr, global row y0 + jr; its
    // colour-1 words: local row 1 + jr, global row y0 - 1 + jr, and (has2)
    // local row 1 + jr2, global row y0 - 1 + jr2
    const uint32_t ya = (y0 + jr) & (L - 1u), yb = (y0 + jr + L - 1u) & (L - 1u);
    const uint32_t yc = (y0 + jr2 + L - 1u) & (L - 1u);
    const uint32_t row0 = (2u + jr) << sw, row1 = (1u + jr) << sw, row2 = (1u + jr2) << sw;
    u32x4 p0{0u, 0u, 0u, 0u}, p0r{0u, 0u, 0u, 0u}, p1{0u, 0u, 0u, 0u}, p1r{0u, 0u, 0u, 0u};
    u32x4 p2{0u, 0u, 0u, 0u}, p2r{0u, 0u, 0u, 0u};  // second words: 256-thread launches only
    if (has0 && n > 0u) metropolis_rng(ya * wpr + q, a.t0, 0u, k0, k1, p0, p0r);
    // neighbour words: local rows 0, 1 from the upper strip's last two rows,
    // R+2, R+3 from the lower strip's first two (one word per thread)
    const bool get_h = jw < hw, get_f = own && jw < fw;
    const uint32_t h_src = jw < 2u * wpr ? up * hw + 2u * wpr + jw : dn * hw + (jw - 2u * wpr);
    const uint32_t h_dst = (jw < 2u * wpr ? 0u : R * wpr) + jw;
    // with split: the samples the measuring workgroup has copied, as last seen
    const uint32_t *snapc = a.ws + sh.o_snapc + w * kStripFlagStride;
    uint64_t *snap = reinterpret_cast<uint64_t *>(a.ws + sh.o_snap);
    uint32_t copied = 0;
    // Device-memory loads are issued before this wave's agent-scope stores
    // and used after them only once other work has passed: a wave's wait
    // for a load also waits for every store it issued before the load's use.
    bool ok = true, bad = false;
    uint32_t vote = 0;
    for (uint32_t it = 0; it < n; ++it) {
        const uint32_t t = a.t0 + it, seq = it + 1u, par = it & 1u;
        const bool pub = own && it >= 2u;
        // the lower strip's first blocked rows of sample it-2, sent during
        // its sweep it-1, loaded before this colour-0 update
        const uint64_t *src_f = xfirst + ((size_t)((it + fs - 2u) % fs) * nwg + dn) * fw + jw;
        uint64_t vf = pub && get_f ? ll_peek(src_f) : 0ull;
        // with split, slot it mod kStripSnapSlots is free once sample
        // it - kStripSnapSlots has been copied
        const uint32_t need = it >= kStripSnapSlots ? it - kStripSnapSlots + 1u : 0u;
        const uint32_t copied_now = sh.split && copied < need ? ld_relaxed(snapc) : copied;
        if (has0) cb0[row0 + q] ^= strip_flip(cb0, cb1, row0, q, wpr, ya, 0u, acc, k0, k1, t, p0, p0r);
        if (pub && get_f) {
            if ((uint32_t)(vf >> 32) != it - 1u) {
                vf = ll_wait(src_f, it - 1u, err, a.strip_sleep);
                if ((uint32_t)(vf >> 32) != it - 1u) bad = true;
            }
            m.dfirst[jw] = (uint32_t)vf;
        }
        // also the barrier between the colour-0 update and the reads of its rows
        if (strip_any(bad, m.red + kRedAny, ++vote)) {
            ok = false;
            break;
        }
        const uint64_t *src_h = xhalo + (size_t)par * nwg * hw + h_src;
        uint64_t vh = get_h ? ll_peek(src_h) : 0ull;
        uint64_t *mine = xhalo + (size_t)(par * nwg + w) * hw;
        for (uint32_t j = threadIdx.x; j < hw; j += NT)
            ll_put(mine + j, cb0[(j < 2u * wpr ? 2u * wpr : (R - 2u) * wpr) + j], seq);
        if (pub && !(ok = strip_publish<NT>(a, m, it - 2u))) break;
        if (own && it >= 1u) {
            strip_measure<NT>(a, m, m.snap + wpr);
            for (uint32_t i = threadIdx.x; i < fw; i += NT)
                ll_put(xfirst + ((size_t)((it - 1u) % fs) * nwg + w) * fw + i, m.frow[i], it);
        }
        if (has1) metropolis_rng(yb * wpr + q, t, 1u, k0, k1, p1, p1r);
        if (NT == 256 && has2) metropolis_rng(yc * wpr + q2, t, 1u, k0, k1, p2, p2r);
        if (has0 && it + 1u < n) metropolis_rng(ya * wpr + q, t + 1u, 0u, k0, k1, p0, p0r);
        if (own && sh.lam > 0u) strip_coins<NT>(a, m, t, k0, k1);
        if (get_h) {
            if ((uint32_t)(vh >> 32) != seq) {
                vh = ll_wait(src_h, seq, err, a.strip_sleep);
                if ((uint32_t)(vh >> 32) != seq) bad = true;
            }
            cb0[h_dst] = (uint32_t)vh;
        }
        if (strip_any(bad, m.red + kRedAny, ++vote)) {
            ok = false;
            break;
        }
        // With split, sample t goes to the measuring workgroup as each word
        // is final: both colours of local rows 1 .. R+2, word i of a colour
        // at slot index i (colour 1 after the first half), written by the
        // thread that updates colour-1 word i.  Colour 0 is final once the
        // neighbours' rows are in; colour 1 goes from the register that
        // computed it.  A failed wait for the slot is voted on at the next
        // barrier.
        uint64_t *dst = nullptr;
        if (sh.split) {
            copied = copied_now;
            if (copied < need) {
                copied = flag_wait(snapc, need, err) ? need : 0u;
                bad = copied < need;
            }
            if (!bad) {
                dst = snap + ((size_t)(it % kStripSnapSlots) * nwg + w) * sh.snap_words;
                if (has1) ll_put(dst + jw, cb0[row1 + q], seq);
                if (NT == 256 && has2) ll_put(dst + j2, cb0[row2 + q2], seq);
            }
        }
        const uint32_t half = sh.snap_words >> 1;
        if (has1) {
            const uint32_t v = cb1[row1 + q] ^ strip_flip(cb1, cb0, row1, q, wpr, yb, 1u, acc, k0, k1, t, p1, p1r);
            cb1[row1 + q] = v;
            if (dst) ll_put(dst + half + jw, v, seq);
        }
        if (NT == 256 && has2) {
            const uint32_t v = cb1[row2 + q2] ^ strip_flip(cb1, cb0, row2, q2, wpr, yc, 1u, acc, k0, k1, t, p2, p2r);
            cb1[row2 + q2] = v;
            if (dst) ll_put(dst + half + j2, v, seq);
        }
        lds_barrier();
        if (own) {
            // sample t, measured during the next sweep's exchange
            for (uint32_t i = threadIdx.x; i < 2u * cbs; i += NT) m.snap[i] = cb0[i];
            lds_barrier();
        }
    }
    if (ok && sh.split && strip_any(bad, m.red + kRedAny, ++vote)) ok = false;
    if (ok && own && n > 0u) {
        if (n >= 2u) {
            if (get_f)
                m.dfirst[jw] = ll_get(xfirst + ((size_t)((n - 2u) % fs) * nwg + dn) * fw + jw, n - 1u, err, bad,
                                      a.strip_sleep);
            ok = !strip_any(bad, m.red + kRedAny, ++vote);
            if (ok) ok = strip_publish<NT>(a, m, n - 2u);
        }
        if (ok) {
            strip_measure<NT>(a, m, m.snap + wpr);
            for (uint32_t i = threadIdx.x; i < fw; i += NT)
                ll_put(xfirst + ((size_t)((n - 1u) % fs) * nwg + w) * fw + i, m.frow[i], n);
            if (get_f)
                m.dfirst[jw] = ll_get(xfirst + ((size_t)((n - 1u) % fs) * nwg + dn) * fw + jw, n, err, bad,
                                      a.strip_sleep);
            ok = !strip_any(bad, m.red + kRedAny, ++vote);
            if (ok) ok = strip_publish<NT>(a, m, n - 1u);
        }
    }
    if (!ok) return;  // the host reports the error word; the lattice is not written back
    for (uint32_t j = threadIdx.x; j < R * wpr; j += NT) {
        const uint32_t r = 2u + (j >> sw), qq = j & mk, y = y0 + (j >> sw);
        const uint32_t ce = y & 1u;
        store64b(lat.spins + (size_t)y * L + 64u * qq, cb0[ce * cbs + r * wpr + qq], cb0[(ce ^ 1u) * cbs + r * wpr + qq]);
    }
    if (w == 0u && threadIdx.x == 0u) {
        lat.state[0] = a.t0 + n;
        lat.state[1] = a.q0 + (sh.measure ? n : 0u);
    }
}

// Reads n (data, seq) words at(i), i = 0 .. n-1, spread over the workgroup:
// four loads in flight per thread before any is waited on; use(i, data) for
// each.  Returns true when a wait gave up.
template <int NT, typename At, typename Use>
__device__ bool ll_gather(uint32_t n, uint32_t seq, uint32_t *err, uint32_t sleep, At at, Use use)
{
    bool bad = false;
    for (uint32_t j = threadIdx.x; j < n; j += 4u * NT) {
        uint64_t v[4];
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) v[k] = j + k * NT < n ? ll_peek(at(j + k * NT)) : 0ull;
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            const uint32_t i = j + k * NT;
            if (i >= n) break;
            if ((uint32_t)(v[k] >> 32) != seq) {
                v[k] = ll_wait(at(i), seq, err, sleep);
                if ((uint32_t)(v[k] >> 32) != seq) bad = true;
            }
            use(i, (uint32_t)v[k]);
        }
    }
    return bad;
}

// The measuring workgroup n_wg + w of a split launch: strip w's samples, one
// per sweep, from the copies the strip publishes (rows y0-1 .. y0+R of both
// colours, into local rows 1 .. R+2 of the halo layout).  Per sample k: the
// tie coins while the copy arrives; the copy, then its slot released to the
// strip; the publication of sample k-1 (the lower measuring workgroup's first
// blocked rows of it were sent one sample ago); the measurement of k and its
// first blocked rows to the upper measuring workgroup.
template <int NT>
__device__ void strip_body_meas(const StripArgs &a, uint32_t *lds)
{
    const StripShape &sh = a.sh;
    const HbmLattice &lat = *a.lat;
    const uint32_t R = sh.R, wpr = sh.wpr, nwg = sh.n_wg;
    const uint32_t w = blockIdx.x - nwg, dn = (w + 1u) & (nwg - 1u);
    const StripLds m = strip_lds(sh, lds, w);
    const uint32_t cbs = m.cbs, fw = sh.first_words, fs = sh.first_slots, n = a.n_sweeps;
    const uint32_t half = sh.snap_words >> 1;
    uint32_t *err = a.ws + sh.o_err;
    uint64_t *xfirst = reinterpret_cast<uint64_t *>(a.ws + sh.o_xfirst);
    const uint64_t *snap = reinterpret_cast<const uint64_t *>(a.ws + sh.o_snap);
    uint32_t *snapc = a.ws + sh.o_snapc + w * kStripFlagStride;
    for (uint32_t i = threadIdx.x; i < kStripRing; i += NT) m.red[i] = 0u;  // consumed counts seen
    for (uint32_t i = threadIdx.x; i < kStripCntWords; i += NT) m.pend[i] = 0u;
    if (threadIdx.x == 0) m.red[kRedAny] = 0u;
    __syncthreads();
    const uint32_t k0 = lat.k0, k1 = lat.k1, jw = threadIdx.x;
    const bool get_f = jw < fw;
    bool ok = true;
    uint32_t vote = 0;
    for (uint32_t k = 0; k < n && ok; ++k) {
        // loads first: sample k-1's first rows from below, sample k's copy
        const uint64_t *src_f = xfirst + ((size_t)((k + fs - 1u) % fs) * nwg + dn) * fw + jw;
        uint64_t vf = k >= 1u && get_f ? ll_peek(src_f) : 0ull;
        const uint64_t *slot = snap + ((size_t)(k % kStripSnapSlots) * nwg + w) * sh.snap_words;
        if (sh.lam > 0u) strip_coins<NT>(a, m, a.t0 + k, k0, k1);
        bool bad = ll_gather<NT>(
            sh.snap_words, k + 1u, err, a.strip_sleep, [&](uint32_t i) { return slot + i; },
            [&](uint32_t i, uint32_t v) {
                const uint32_t c = i >= half ? 1u : 0u;
                m.cb0[c * cbs + wpr + (i - c * half)] = v;
            });
        if (k >= 1u && get_f) {
            if ((uint32_t)(vf >> 32) != k) {
                vf = ll_wait(src_f, k, err, a.strip_sleep);
                if ((uint32_t)(vf >> 32) != k) bad = true;
            }
            m.dfirst[jw] = (uint32_t)vf;
        }
        if (strip_any(bad, m.red + kRedAny, ++vote)) {
            ok = false;
            break;
        }
        // every thread's copy of the slot is in LDS: the strip may reuse it
        if (threadIdx.x == 0)
            __hip_atomic_store(snapc, k + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k >= 1u && !(ok = strip_publish<NT>(a, m, k - 1u))) break;
        strip_measure<NT>(a, m, m.cb0 + wpr);
        for (uint32_t i = threadIdx.x; i < fw; i += NT)
            ll_put(xfirst + ((size_t)(k % fs) * nwg + w) * fw + i, m.frow[i], k + 1u);
    }
    if (!ok || n == 0u) return;
    // the last sample: its first rows from below, then its publication
    bool bad = false;
    if (get_f)
        m.dfirst[jw] = ll_get(xfirst + ((size_t)((n - 1u) % fs) * nwg + dn) * fw + jw, n, err, bad, a.strip_sleep);
    if (!strip_any(bad, m.red + kRedAny, ++vote)) (void)strip_publish<NT>(a, m, n - 1u);
}

// A tail workgroup: samples q = tail, tail + n_tail, ...  For each, once every
// strip has published it: the summed counts of levels 0..lam, the levels
// beyond lam from the gathered level-lam rows (one per strip), the operator
// values and the record terms.
template <int NT>
__device__ void tail_body(const StripArgs &a, uint32_t *lds)
{
    const StripShape &sh = a.sh;
    const HbmLattice &lat = *a.lat;
    const uint32_t L = sh.L, nl = sh.nl, lam = sh.lam, nwg = sh.n_wg, pw = sh.per_wg, cw = sh.cnt_words;
    const uint32_t k0 = lat.k0, k1 = lat.k1;
    int64_t *S = reinterpret_cast<int64_t *>(lds);  // [kMaxLevels + 1][3], then M
    uint32_t *cnt = lds + 2u * 3u * (kMaxLevels + 1u) + 2u;
    uint32_t *red = cnt + kStripCntWords;
    uint32_t *A = red + kRedWords;
    const uint32_t side = L >> lam;
    uint32_t *B = A + (nl > lam ? (side * side) / 32u : 0u);
    uint32_t *ws = a.ws;
    uint32_t *err = ws + sh.o_err;
    const uint32_t rs = record_size(nl);
    for (uint32_t q = blockIdx.x - nwg - sh.n_meas; q < a.n_sweeps; q += sh.n_tail) {
        const uint32_t s = q % kStripRing, gen = q / kStripRing, t = a.t0 + q, seq = q + 1u;
        const uint64_t *slot = reinterpret_cast<const uint64_t *>(ws + sh.o_ring) + (size_t)s * nwg * pw;
        for (uint32_t i = threadIdx.x; i < kStripCntWords; i += NT) cnt[i] = 0u;
        // one thread waits for a word of the last strip before the whole
        // workgroup polls: the strips publish a sample at about the same time
        if (threadIdx.x == 0) {
            bool b = false;
            (void)ll_get(slot + (size_t)(nwg - 1u) * pw, seq, err, b, a.tail_sleep);
        }
        __syncthreads();
        // every strip's counts of levels 0..lam, and its level-lam row
        const bool bad = ll_gather<NT>(
            nwg * pw, seq, err, a.tail_sleep, [&](uint32_t i) { return slot + i; },
            [&](uint32_t i, uint32_t v) {
                const uint32_t k = i % pw;
                if (k < cw) {
                    if (v) atomicAdd(cnt + k, v);
                } else {
                    A[(i / pw) * (pw - cw) + (k - cw)] = v;
                }
            });
        if (__syncthreads_or(bad ? 1 : 0)) return;
        if (nl > lam) {
            uint32_t *cur = A;
            uint32_t Ll = side;
            for (uint32_t l = lam + 1u; l <= nl; ++l) {
                uint32_t *dst = ((l - lam) & 1u) ? B : A;
                for (uint32_t i = threadIdx.x; i < rm_block_items(Ll); i += NT)
                    dst[i] = rm_block_item(cur, Ll, i, tie_coin_word(i, t, l - 1u, k0, k1));
                __syncthreads();
                cur = dst;
                Ll >>= 1;
                uint32_t d0 = 0, d1 = 0, d2 = 0;
                for (uint32_t i = threadIdx.x; i < rm_ops_items(Ll); i += NT) rm_ops_item(cur, Ll, i, d0, d1, d2);
                wave_add(cnt + 4u * l, 3u, d0, d1, d2, 0u);
            }
        }
        __syncthreads();
        // operator values and the sample's record terms
        for (uint32_t i = threadIdx.x; i <= 3u * (nl + 1u); i += NT) {
            if (i == 3u * (nl + 1u)) {
                S[3u * (kMaxLevels + 1u)] = (int64_t)L * L - 2 * (int64_t)cnt[3];
                continue;
            }
            const uint32_t l = i / 3u, k = i % 3u;
            const int64_t n = (int64_t)(L >> l) * (L >> l);
            S[i] = (k == 2u ? n : 2 * n) - 2 * (int64_t)cnt[4u * l + k];
        }
        __syncthreads();
        const int64_t(*Sv)[3] = reinterpret_cast<const int64_t(*)[3]>(S);
        const int64_t M = S[3u * (kMaxLevels + 1u)];
        int64_t *rec = lat.records + (size_t)((a.q0 + q) / (uint32_t)lat.sweeps_per_bin) * rs;
        for (uint32_t j = threadIdx.x; j < rs; j += NT)
            atomicAdd(reinterpret_cast<unsigned long long *>(rec + j), (unsigned long long)record_term(j, nl, Sv, M));
        // every read of the slot is complete (its values were used above)
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(ws + sh.o_consumed + s * kStripFlagStride, gen + 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_strip(StripArgs a)
{
    extern __shared__ uint32_t lds[];
    if (blockIdx.x >= a.sh.n_wg + a.sh.n_meas) tail_body<NT>(a, lds);
    else if (blockIdx.x >= a.sh.n_wg) strip_body_meas<NT>(a, lds);
    else if (a.sh.halo) strip_body_halo<NT>(a, lds);
    else strip_body<NT>(a, lds);
}

template <int NT>
static hipError_t launch_nt(const StripArgs &a, hipStream_t s)
{
    const uint32_t lds_bytes = a.sh.lds_words * 4u;
    static std::mutex mu;
    static uint32_t opted_in = 65536u;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (lds_bytes > opted_in) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_strip<NT>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
            if (e != hipSuccess) return e;
            opted_in = lds_bytes;
        }
    }
    StripArgs arg = a;
    const dim3 grid(a.sh.n_wg + a.sh.n_meas + a.sh.n_tail);
    // MCRG_STRIP_NONCOOP=1: an ordinary launch, for profilers that fail on
    // cooperative ones.  The runner has checked that the grid fits the device
    // at once, and on an idle device every workgroup is then dispatched; if
    // one were not, the strips' bounded waits would report an error.
    static const bool noncoop = [] {
        const char *v = std::getenv("MCRG_STRIP_NONCOOP");
        return v && std::atoi(v) != 0;
    }();
    if (noncoop) {
        hipLaunchKernelGGL(k_strip<NT>, grid, dim3(NT), lds_bytes, s, arg);
        return hipGetLastError();
    }
    void *args[] = {&arg};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_strip<NT>), grid, dim3(NT), args, lds_bytes,
                                      s);
}

hipError_t launch_strip(const StripArgs &a, int threads, hipStream_t s)
{
    const uint32_t blocks = std::min<uint32_t>(1024u, (a.sh.ws_words + 255u) / 256u);
    hipLaunchKernelGGL(k_strip_reset, dim3(blocks), dim3(256), 0, s, a.ws, a.sh.ws_words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (threads) {
    case 64: return launch_nt<64>(a, s);
    case 256: return launch_nt<256>(a, s);
    case 512: return launch_nt<512>(a, s);
    case 1024: return launch_nt<1024>(a, s);
    default: return hipErrorInvalidValue;
    }
}

int strip_max_resident(int threads, uint32_t lds_bytes)
{
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (threads) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_strip<64>, 64, lds_bytes); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_strip<256>, 256, lds_bytes); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_strip<512>, 512, lds_bytes); break;
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_strip<1024>, 1024, lds_bytes); break;
    default: break;
    }
    if (e != hipSuccess || per_cu <= 0) return 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace mcrg
