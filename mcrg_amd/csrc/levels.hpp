// Measurement on bit-packed lattices, one work item at a time so that a caller
// can spread items over the threads of a workgroup (and over many lattices).
//
// RM layout ("row-major bits"): site (x, y) of an Ll x Ll lattice is bit
// m = y*Ll + x, word m >> 5; rows shorter than 32 bits share words.  Levels
// l >= 1 of every schedule use it.
//
// CB layout ("checkerboard bits"): the L^2/2 sites of colour c in canonical
// order m = y*(L/2) + k (x = 2k + ((y + c) & 1)), word m >> 5; colour 0 words
// then colour 1 words.  Level 0 of the LDS-resident schedule uses it (L >= 8).
//
// Operators are returned as disagreement counts (see hbm_kernels.hip):
// NN = 2n - 2 dNN, NNN = 2n - 2 dNNN, P = n - 2 dP for an n-site level.
#pragma once

#include "device_common.hpp"

namespace mcrg {

// log2 of a power of two
__device__ __forceinline__ uint32_t ilog2u(uint32_t v) { return (uint32_t)__builtin_ctz(v); }

// ------------------------------------------------------------------ RM levels
__device__ __forceinline__ uint32_t rm_row_small(const uint32_t *buf, uint32_t Ll, uint32_t y)
{
    const uint32_t m = y * Ll;
    const uint32_t mask = Ll >= 32u ? 0xFFFFFFFFu : ((1u << Ll) - 1u);
    return (buf[m >> 5] >> (m & 31u)) & mask;
}

__device__ __forceinline__ uint32_t rot_right_small(uint32_t v, uint32_t Ll)
{
    // bit x of the result = bit x+1 (mod Ll) of v: the row of right neighbours.
    return ((v >> 1) | ((v & 1u) << (Ll - 1u))) & ((1u << Ll) - 1u);
}

// Work items of rm_ops_item: one per word (Ll >= 32) or per row (Ll < 32).
__device__ __forceinline__ uint32_t rm_ops_items(uint32_t Ll) { return Ll >= 32u ? Ll * (Ll >> 5) : Ll; }

__device__ __forceinline__ void rm_ops_item(const uint32_t *buf, uint32_t Ll, uint32_t i, uint32_t &dnn,
                                            uint32_t &dnnn, uint32_t &dp)
{
    uint32_t a, r, d, g;
    if (Ll >= 32u) {
        const uint32_t wpr = Ll >> 5, sh = ilog2u(wpr);
        const uint32_t y = i >> sh, q = i & (wpr - 1u);
        const uint32_t yn = (y + 1u) & (Ll - 1u), qn = (q + 1u) & (wpr - 1u);
        a = buf[y * wpr + q];
        const uint32_t an = buf[y * wpr + qn];
        d = buf[yn * wpr + q];
        const uint32_t dn = buf[yn * wpr + qn];
        r = (a >> 1) | (an << 31);
        g = (d >> 1) | (dn << 31);
    } else {
        a = rm_row_small(buf, Ll, i);
        d = rm_row_small(buf, Ll, (i + 1u) & (Ll - 1u));
        r = rot_right_small(a, Ll);
        g = rot_right_small(d, Ll);
    }
    dnn += (uint32_t)(popc(a ^ r) + popc(a ^ d));
    dnnn += (uint32_t)(popc(a ^ g) + popc(r ^ d));
    dp += (uint32_t)popc(xor3(a, r, d) ^ g);
}

// Output words of rm_block_item (level of side Ll -> Ll/2).
__device__ __forceinline__ uint32_t rm_block_items(uint32_t Ll)
{
    const uint32_t Lb = Ll >> 1;
    return (Lb * Lb + 31u) >> 5;
}

// Output word i of the b = 2 majority blocking of an RM level; ties take bits
// of coin word i (the coin word index equals the output word index).
__device__ __forceinline__ uint32_t rm_block_item(const uint32_t *src, uint32_t Ll, uint32_t i,
                                                  uint32_t coin)
{
    const uint32_t Lb = Ll >> 1;
    uint32_t out = 0;
    if (Lb >= 32u) {
        const uint32_t wpo = Lb >> 5, wpi = Ll >> 5;
        const uint32_t by = i >> ilog2u(wpo), q = i & (wpo - 1u);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t w0 = src[(2u * by) * wpi + 2u * q + (uint32_t)h];
            const uint32_t w1 = src[(2u * by + 1u) * wpi + 2u * q + (uint32_t)h];
            out |= block16(w0, w1, (coin >> (16 * h)) & 0xFFFFu) << (16 * h);
        }
    } else {
        // rows of Lb < 32 bits: this word holds blocked rows by0 .. by0 + 32/Lb - 1
        const uint32_t lb = ilog2u(Lb);
        const uint32_t rows = (32u >> lb) < Lb ? (32u >> lb) : Lb;
        const uint32_t by0 = (32u * i) >> lb;
        const uint32_t mask = (1u << Lb) - 1u;
        for (uint32_t rr = 0; rr < rows; ++rr) {
            const uint32_t by = by0 + rr;
            if (by >= Lb) break;
            const uint32_t w0 = rm_row_small(src, Ll, 2u * by), w1 = rm_row_small(src, Ll, 2u * by + 1u);
            const uint32_t off = (by * Lb) & 31u;
            out |= (block16(w0, w1, coin >> off) & mask) << off;
        }
    }
    return out;
}

// ------------------------------------------------------------------ CB level 0
// Segment masks for L in {8, 16, 32}: a word holds 64/L rows of r = L/2 bits.
struct SegMasks {
    uint32_t r;     // segment width
    uint32_t low;   // bit 0 of every segment
    uint32_t high;  // bit r-1 of every segment
    uint32_t even;  // all bits of the even segments
};

__device__ __forceinline__ SegMasks seg_masks(uint32_t L)
{
    SegMasks m;
    m.r = L >> 1;
    m.low = m.r == 4u ? 0x11111111u : m.r == 8u ? 0x01010101u : 0x00010001u;
    m.high = m.low << (m.r - 1u);
    m.even = m.r == 4u ? 0x0F0F0F0Fu : m.r == 8u ? 0x00FF00FFu : 0x0000FFFFu;
    return m;
}

__device__ __forceinline__ uint32_t rotl_seg(uint32_t v, const SegMasks &m)
{
    return ((v << 1) & ~m.low) | ((v >> (m.r - 1u)) & m.low);
}

__device__ __forceinline__ uint32_t rotr_seg(uint32_t v, const SegMasks &m)
{
    return ((v >> 1) & ~m.high) | ((v << (m.r - 1u)) & m.high);
}

// Keeps the even r-bit segments and packs them into the low 16 bits.
__device__ __forceinline__ uint32_t compress_seg(uint32_t x, uint32_t r)
{
    if (r == 4u) {
        x &= 0x0F0F0F0Fu;
        x = (x | (x >> 4)) & 0x00FF00FFu;
        return (x | (x >> 8)) & 0x0000FFFFu;
    }
    if (r == 8u) {
        x &= 0x00FF00FFu;
        return (x | (x >> 8)) & 0x0000FFFFu;
    }
    return x & 0x0000FFFFu;
}

// The four nearest-neighbour words of CB word w of colour c (the neighbours
// are the other colour, `oth`).  W = words per colour, L >= 8.
struct Nbr4 {
    uint32_t nu, nd, nl, nr;
};

__device__ __forceinline__ Nbr4 cb_neighbours(const uint32_t *oth, uint32_t L, uint32_t W, uint32_t w,
                                              uint32_t c)
{
    Nbr4 n;
    if (L >= 64u) {
        const uint32_t wpr = L >> 6, sh = ilog2u(wpr);
        const uint32_t y = w >> sh, q = w & (wpr - 1u);
        const uint32_t p = (y + c) & 1u;
        const uint32_t row = y << sh;
        // the word beside o that holds the row's wrap-around neighbour: word
        // q - 1 on p = 0 rows, q + 1 on p = 1 rows (one load either way, so
        // that every load of the word issues before one wait)
        const uint32_t o = oth[row + q];
        const uint32_t side = oth[row + ((q + 2u * p - 1u) & (wpr - 1u))];
        n.nu = oth[(((y - 1u) & (L - 1u)) << sh) + q];
        n.nd = oth[(((y + 1u) & (L - 1u)) << sh) + q];
        n.nr = p == 0u ? o : (o >> 1) | (side << 31);
        n.nl = p == 0u ? (o << 1) | (side >> 31) : o;
    } else {
        const SegMasks m = seg_masks(L);
        const uint32_t o = oth[w];
        const uint32_t op = oth[(w - 1u) & (W - 1u)], on = oth[(w + 1u) & (W - 1u)];
        n.nu = (o << m.r) | (op >> (32u - m.r));
        n.nd = (o >> m.r) | (on << (32u - m.r));
        const uint32_t m0 = c == 0u ? m.even : ~m.even;  // segments whose row has p = 0
        n.nl = (rotl_seg(o, m) & m0) | (o & ~m0);
        n.nr = (o & m0) | (rotr_seg(o, m) & ~m0);
    }
    return n;
}

// Level-0 operator counts contributed by CB word w of colour c: NN from the
// 4 bonds of colour-0 sites (every bond has one colour-0 end), NNN from the 2
// down-diagonal bonds of every site, P from the plaquette whose top-left
// corner is the site, ndown = number of -1 spins.
__device__ __forceinline__ void cb_ops_item(const uint32_t *own, const uint32_t *oth, uint32_t L,
                                            uint32_t W, uint32_t w, uint32_t c, uint32_t &dnn,
                                            uint32_t &dnnn, uint32_t &dp, uint32_t &ndown)
{
    const uint32_t s = own[w];
    const Nbr4 n = cb_neighbours(oth, L, W, w, c);
    uint32_t dr, dl;  // same colour, row y+1: (x+1, y+1) and (x-1, y+1)
    if (L >= 64u) {
        const uint32_t wpr = L >> 6, sh = ilog2u(wpr);
        const uint32_t y = w >> sh, q = w & (wpr - 1u);
        const uint32_t p = (y + c) & 1u;
        const uint32_t row = ((y + 1u) & (L - 1u)) << sh;
        const uint32_t nx = own[row + q];
        const uint32_t side = own[row + ((q + 2u * p - 1u) & (wpr - 1u))];
        dr = p == 0u ? nx : (nx >> 1) | (side << 31);
        dl = p == 0u ? (nx << 1) | (side >> 31) : nx;
    } else {
        const SegMasks m = seg_masks(L);
        const uint32_t sn = own[(w + 1u) & (W - 1u)];
        const uint32_t nx = (s >> m.r) | (sn << (32u - m.r));
        const uint32_t m0 = c == 0u ? m.even : ~m.even;
        dr = (nx & m0) | (rotr_seg(nx, m) & ~m0);
        dl = (rotl_seg(nx, m) & m0) | (nx & ~m0);
    }
    if (c == 0u)
        dnn += (uint32_t)(popc(s ^ n.nu) + popc(s ^ n.nd) + popc(s ^ n.nl) + popc(s ^ n.nr));
    dnnn += (uint32_t)(popc(s ^ dr) + popc(s ^ dl));
    dp += (uint32_t)popc(xor3(s, n.nr, n.nd) ^ dr);
    ndown += (uint32_t)popc(s);
}

// Work items of cb_block_item: level-1 RM words.
__device__ __forceinline__ uint32_t cb_block_items(uint32_t L) { return (L * L / 4u + 31u) >> 5; }

// Level-1 RM word i from the CB level 0 (c0, c1 = colour 0 and 1 words); ties
// take bits of coin word i.  The four spins of block (bx, by) sit at the same
// bit, k = bx, of the colour-0 and colour-1 words of rows 2by and 2by+1.
__device__ __forceinline__ uint32_t cb_block_item(const uint32_t *c0, const uint32_t *c1, uint32_t L,
                                                  uint32_t W, uint32_t i, uint32_t coin)
{
    if (L >= 64u) {
        const uint32_t wpr = L >> 6, sh = ilog2u(wpr);
        const uint32_t by = i >> sh, q = i & (wpr - 1u);
        const uint32_t r0 = ((2u * by) << sh) + q, r1 = ((2u * by + 1u) << sh) + q;
        const uint32_t b1 = c0[r0], b2 = c1[r0], b3 = c0[r1], b4 = c1[r1];
        const uint32_t maj3 = (b1 & b2 & (b3 | b4)) | (b3 & b4 & (b1 | b2));
        const uint32_t tie = ~(xor3(b1, b2, b3) ^ b4) & ~(b1 & b2 & b3 & b4) & (b1 | b2 | b3 | b4);
        return maj3 | (tie & coin);
    }
    // L in {8, 16, 32}: CB word w holds rows w*R .. w*R + R - 1 (R = 64/L, even),
    // i.e. R/2 blocked rows = 16 output bits; output word i takes words 2i, 2i+1.
    const SegMasks m = seg_masks(L);
    uint32_t out = 0;
    for (uint32_t h = 0; h < 2u; ++h) {
        const uint32_t w = 2u * i + h;
        if (w >= W) break;
        const uint32_t a0 = c0[w], a1 = c1[w];
        const uint32_t b1 = a0 & m.even, b2 = a1 & m.even;
        const uint32_t b3 = (a0 >> m.r) & m.even, b4 = (a1 >> m.r) & m.even;
        const uint32_t maj3 = (b1 & b2 & (b3 | b4)) | (b3 & b4 & (b1 | b2));
        const uint32_t tie = ~(xor3(b1, b2, b3) ^ b4) & ~(b1 & b2 & b3 & b4) & (b1 | b2 | b3 | b4) & m.even;
        const uint32_t c16 = (coin >> (16u * h)) & 0xFFFFu;
        out |= (compress_seg(maj3, m.r) | (compress_seg(tie, m.r) & c16)) << (16u * h);
    }
    return out;
}

}  // namespace mcrg
