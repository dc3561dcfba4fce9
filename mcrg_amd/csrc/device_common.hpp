// Device-side building blocks shared by every MCRG kernel (gfx950, wave64).
//
// Stream layout (identical for every schedule and storage layout, so that any
// schedule reproduces the same chain bit for bit):
//   key      (lo32(seed), hi32(seed))
//   sweep    Philox(x0 = word, x1 = t, x2 = 0, x3 = 1<<28 | colour<<24): out[q] is
//            bit 31-q of every site's 32-bit uniform; sites whose 4 top bits tie
//            with the threshold's take the top 28 bits of word rho & 3 of
//            Philox(word, t, 1 + (rho >> 2), 1<<28 | colour<<24), rho = their rank
//   tie coin Philox(x0 = bm>>7, x1 = t, 0, 2<<28 | level<<16); bit bm&31 of out[(bm>>5)&3]
//   init     Philox(x0 = m>>7, 0, 0, 3<<28);                  bit m&31 of out[(m>>5)&3]
// Site bits: b = 1 for spin -1 (the sign bit of the int8 value), b = 0 for +1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcrg {

constexpr uint32_t kTagSweep = 1u;
constexpr uint32_t kTagTie = 2u;
constexpr uint32_t kTagInit = 3u;
constexpr uint32_t kTagSwBond = 4u;
constexpr uint32_t kTagSwFlip = 5u;

struct u32x4 {
    uint32_t v0, v1, v2, v3;
};

// a ^ b ^ c as one v_bitop3_b32 (gfx950); truth table 0x96 = odd parity of the
// three inputs.  The compiler emits two v_xor_b32 for the plain expression.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (Salmon et al., SC'11).  Ten rounds of two 32x32->64 multiplies;
// on CDNA each multiply is one v_mad_u64_u32, each round's key mix one xor3.
__device__ __forceinline__ u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t pick4(const u32x4 &r, uint32_t q)
{
    return q == 0 ? r.v0 : q == 1 ? r.v1 : q == 2 ? r.v2 : r.v3;
}

// Acceptance parameters of one chain.  thr[q] is floor(exp(-4J(q+1)) 2^32) for
// the sites whose e = s*h is 2 (q = 0) or 4 (q = 1); always[q] = 1 when that
// value is 2^32 (J = 0).
struct Accept {
    uint32_t thr0, thr1;
    uint32_t always0, always1;
};

// Metropolis decision for the 32 sites of one word.  s, nu, nd, nl, nr are the
// spin bits of the sites and of their up/down/left/right neighbours; `valid`
// masks the sites that exist.  Returns the bits to flip.
//
// e <= 0 flips outright.  The other sites compare a 32-bit uniform U_i with the
// threshold of their class.  The first Philox call gives every site the 4 top
// bits of its U_i (slice q = out[q] supplies bit 31-q of all 32 sites), which
// decide the comparison unless they equal the threshold's 4 top bits
// (probability 1/16).  The sites still open, ranked by bit index, take their 28
// low bits from refinement calls, four sites per call.
__device__ __forceinline__ uint32_t metropolis_word(uint32_t s, uint32_t nu, uint32_t nd,
                                                    uint32_t nl, uint32_t nr, uint32_t valid,
                                                    const Accept &acc, uint32_t k0, uint32_t k1,
                                                    uint32_t word, uint32_t t, uint32_t colour)
{
    const uint32_t a1 = s ^ nu, a2 = s ^ nd, a3 = s ^ nl, a4 = s ^ nr;
    const uint32_t x1 = a1 ^ a2, y1 = a1 & a2, x2 = a3 ^ a4, y2 = a3 & a4;
    const uint32_t any = a1 | a2 | a3 | a4;
    const uint32_t c1 = (x1 ^ x2) & ~(y1 | y2);  // one anti-aligned neighbour: e = 2
    const uint32_t c2 = ~any;                    // none: e = 4
    uint32_t flip = any & ~c1;                   // two or more: e <= 0
    flip |= (acc.always0 ? c1 : 0u) | (acc.always1 ? c2 : 0u);
    const uint32_t m1 = acc.always0 ? 0u : c1;
    const uint32_t m2 = acc.always1 ? 0u : c2;
    uint32_t open = (m1 | m2) & valid;
    uint32_t accept = 0;
    if (open != 0u) {
        const uint32_t tag = (kTagSweep << 28) | (colour << 24);
        // the first refinement call is drawn beside the first-stage call: a
        // wave nearly always needs it, and the two independent Philox chains
        // interleave
        const u32x4 r = philox(word, t, 0u, tag, k0, k1);
        u32x4 ref = philox(word, t, 1u, tag, k0, k1);
        const uint32_t rq[4] = {r.v0, r.v1, r.v2, r.v3};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t b0 = 0u - ((acc.thr0 >> (31 - q)) & 1u);
            const uint32_t b1 = 0u - ((acc.thr1 >> (31 - q)) & 1u);
            const uint32_t tb = (m1 & b0) | (m2 & b1);
            accept |= open & ~rq[q] & tb;
            open &= ~(rq[q] ^ tb);
        }
        for (uint32_t rank = 0; open != 0u; ++rank) {
            const uint32_t i = (uint32_t)__builtin_ctz(open);
            open &= open - 1u;
            if (rank != 0u && (rank & 3u) == 0u) ref = philox(word, t, 1u + (rank >> 2), tag, k0, k1);
            const uint32_t low = pick4(ref, rank & 3u) >> 4;
            const uint32_t thr = ((m1 >> i) & 1u) ? acc.thr0 : acc.thr1;
            accept |= (low < (thr & 0x0FFFFFFFu) ? 1u : 0u) << i;
        }
    }
    return (flip | accept) & valid;
}

// The two Philox calls metropolis_word draws first for a word (the 4 top bits
// of every site's uniform, and the first refinement call), which depend only
// on the word, sweep and colour: a caller can draw them ahead of the spins.
__device__ __forceinline__ void metropolis_rng(uint32_t word, uint32_t t, uint32_t colour, uint32_t k0, uint32_t k1,
                                               u32x4 &r, u32x4 &ref)
{
    const uint32_t tag = (kTagSweep << 28) | (colour << 24);
    r = philox(word, t, 0u, tag, k0, k1);
    ref = philox(word, t, 1u, tag, k0, k1);
}

// metropolis_word with the first two Philox calls drawn by metropolis_rng
// (the same decisions bit for bit).
__device__ __forceinline__ uint32_t metropolis_word_drawn(uint32_t s, uint32_t nu, uint32_t nd, uint32_t nl,
                                                          uint32_t nr, const Accept &acc, uint32_t k0, uint32_t k1,
                                                          uint32_t word, uint32_t t, uint32_t colour,
                                                          const u32x4 &r, u32x4 ref)
{
    const uint32_t a1 = s ^ nu, a2 = s ^ nd, a3 = s ^ nl, a4 = s ^ nr;
    const uint32_t x1 = a1 ^ a2, y1 = a1 & a2, x2 = a3 ^ a4, y2 = a3 & a4;
    const uint32_t any = a1 | a2 | a3 | a4;
    const uint32_t c1 = (x1 ^ x2) & ~(y1 | y2);
    const uint32_t c2 = ~any;
    uint32_t flip = any & ~c1;
    flip |= (acc.always0 ? c1 : 0u) | (acc.always1 ? c2 : 0u);
    const uint32_t m1 = acc.always0 ? 0u : c1;
    const uint32_t m2 = acc.always1 ? 0u : c2;
    uint32_t open = m1 | m2;
    uint32_t accept = 0;
    const uint32_t tag = (kTagSweep << 28) | (colour << 24);
    const uint32_t rq[4] = {r.v0, r.v1, r.v2, r.v3};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t b0 = 0u - ((acc.thr0 >> (31 - q)) & 1u);
        const uint32_t b1 = 0u - ((acc.thr1 >> (31 - q)) & 1u);
        const uint32_t tb = (m1 & b0) | (m2 & b1);
        accept |= open & ~rq[q] & tb;
        open &= ~(rq[q] ^ tb);
    }
    for (uint32_t rank = 0; open != 0u; ++rank) {
        const uint32_t i = (uint32_t)__builtin_ctz(open);
        open &= open - 1u;
        if (rank != 0u && (rank & 3u) == 0u) ref = philox(word, t, 1u + (rank >> 2), tag, k0, k1);
        const uint32_t low = pick4(ref, rank & 3u) >> 4;
        const uint32_t thr = ((m1 >> i) & 1u) ? acc.thr0 : acc.thr1;
        accept |= (low < (thr & 0x0FFFFFFFu) ? 1u : 0u) << i;
    }
    return flip | accept;
}

// Spin bits of the 32 bytes at even (p = 0) or odd (p = 1) positions of a 64-byte
// chunk held as 16 dwords: result bit j = sign bit of byte 2j + p.
__device__ __forceinline__ uint32_t gather_parity(const uint32_t d[16], uint32_t p)
{
    uint32_t out = 0;
    const uint32_t sh0 = 8u * p + 7u, sh1 = 8u * p + 23u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        out |= ((d[i] >> sh0) & 1u) << (2 * i);
        out |= ((d[i] >> sh1) & 1u) << (2 * i + 1);
    }
    return out;
}

// Sign bits of 4 bytes -> 4 contiguous bits.
__device__ __forceinline__ uint32_t sign4(uint32_t d)
{
    return ((d >> 7) & 1u) | ((d >> 14) & 2u) | ((d >> 21) & 4u) | ((d >> 28) & 8u);
}

// Keeps the bits at even positions of x and packs them into the low 16 bits.
__device__ __forceinline__ uint32_t compress_even(uint32_t x)
{
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

// b = 2 majority blocks of a pair of row chunks: block i covers bits 2i, 2i+1 of
// w0 (row 2by) and of w1 (row 2by+1).  Returns the 16 blocked spin bits; a 2-2
// tie takes bit i of coin16.
__device__ __forceinline__ uint32_t block16(uint32_t w0, uint32_t w1, uint32_t coin16)
{
    const uint32_t E = 0x55555555u;
    const uint32_t b1 = w0 & E, b2 = (w0 >> 1) & E, b3 = w1 & E, b4 = (w1 >> 1) & E;
    const uint32_t maj3 = (b1 & b2 & (b3 | b4)) | (b3 & b4 & (b1 | b2));
    const uint32_t par = xor3(b1, b2, b3) ^ b4;
    const uint32_t all = b1 & b2 & b3 & b4;
    const uint32_t some = b1 | b2 | b3 | b4;
    const uint32_t tie = ~par & ~all & some & E;
    return compress_even(maj3) | (compress_even(tie) & coin16);
}

__device__ __forceinline__ uint32_t tie_coin_word(uint32_t cw, uint32_t t, uint32_t level,
                                                  uint32_t k0, uint32_t k1)
{
    const u32x4 r = philox(cw >> 2, t, 0u, (kTagTie << 28) | (level << 16), k0, k1);
    return pick4(r, cw & 3u);
}

__device__ __forceinline__ int popc(uint32_t x) { return __popc(x); }

}  // namespace mcrg
