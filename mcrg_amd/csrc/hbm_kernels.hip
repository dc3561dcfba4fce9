// HBM-resident schedule: int8 lattices in global memory, one launch per
// half-sweep, then two launches per measurement.  Used for the plain
// single-lattice path and for batches of large lattices (one grid row per
// lattice, blockIdx.y).
//
//   k_int8_init           Philox +-1 start        (Lattice::initialize_random_spins, src/lattice.cpp:33-41)
//   k_int8_half_sweep     checkerboard Metropolis (replaces Wolff, src/ising.cpp:87-149)
//   k_sw_local, k_sw_boundary, k_sw_flip
//                         Swendsen-Wang cluster sweep: the Wolff construction
//                         (src/ising.cpp:96-149) for every cluster at once, with
//                         GPU union-find on atomicMin

//   k_int8_measure0       level-0 NN/NNN/P + M and b=2 majority blocking into a
//                         bit-packed level-1 lattice (src/lattice.cpp:102-120, src/mcrg.cpp:314-348)
//   k_rm_level            one large blocked level: operators and the next
//                         blocking, over a grid (src/mcrg.cpp:80-86, src/lattice.cpp:84-120)
//   k_levels_accumulate   the small levels in one workgroup per lattice, then
//                         the per-sample sums and products into the bin record
//                         (src/mcrg.cpp:72-98)
#include "device_common.hpp"
#include "hbm_kernels.hpp"
#include "levels.hpp"

namespace mcrg {

// ---------------------------------------------------------------- init
__global__ __launch_bounds__(256) void k_int8_init(const HbmLattice *__restrict__ lats)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L;
    const uint32_t n = L * L;
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const u32x4 r = philox(m >> 7, 0u, 0u, kTagInit << 28, lat.k0, lat.k1);
    const uint32_t bit = (pick4(r, (m >> 5) & 3u) >> (m & 31u)) & 1u;
    lat.spins[m] = (int8_t)(1 - 2 * (int)bit);
}

// ---------------------------------------------------------------- Swendsen-Wang
// Cluster labels form a union-find forest in which every site points to a
// smaller (or equal) site index, so each tree's root is the smallest site of
// its cluster whatever the order of the unions.  Two passes build it:
//   k_sw_local     one workgroup per tile of up to 32 x 32 sites: bond draws,
//                  union-find of the bonds inside the tile in LDS, then every
//                  site's global label = its tile-local root (tile rows are
//                  row-major segments, so the local minimum is the global one)
//   k_sw_boundary  bonds that leave a tile (right edge, bottom edge), united
//                  in the global labels with atomicMin
// Loads of global labels may return stale values while other threads unite:
// a label only decreases and always names a site of the same cluster, so a
// stale read lengthens a path but never leaves the cluster, and the atomics
// decide.
constexpr uint32_t kSwTile = 32;

__device__ __forceinline__ int32_t sw_root(const int32_t *labels, int32_t i)
{
    for (;;) {
        const int32_t p = __hip_atomic_load(labels + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == i) return i;
        i = p;
    }
}

__device__ __forceinline__ void sw_unite(int32_t *labels, int32_t a, int32_t b)
{
    for (;;) {
        a = sw_root(labels, a);
        b = sw_root(labels, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        // link root b below a; if b stopped being a root, continue from its parent
        const int32_t old = atomicMin(labels + b, a);
        if (old == b) return;
        b = old;
    }
}

__device__ __forceinline__ int32_t lds_root(const int32_t *lab, int32_t i)
{
    for (;;) {
        const int32_t p = lab[i];
        if (p == i) return i;
        i = p;
    }
}

__device__ __forceinline__ void lds_unite(int32_t *lab, int32_t a, int32_t b)
{
    for (;;) {
        a = lds_root(lab, a);
        b = lds_root(lab, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicMin(lab + b, a);
        if (old == b) return;
        b = old;
    }
}

// Tile of TW x TH sites, TW = TH = min(L, 32); blockIdx.x = tile (row-major).
// Thread pairs of horizontally adjacent sites (2p, 2p+1) share one Philox call,
// which holds both sites' right and down bond draws.
__global__ __launch_bounds__(256) void k_sw_local(const HbmLattice *__restrict__ lats, uint32_t t_off)
{
    __shared__ int32_t lab[kSwTile * kSwTile];
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L, mask = L - 1u;
    const uint32_t TW = L < kSwTile ? L : kSwTile, nt = TW * TW;
    const uint32_t tiles_x = L / TW;
    const uint32_t x0 = (blockIdx.x % tiles_x) * TW, y0 = (blockIdx.x / tiles_x) * TW;
    const bool wrap = TW == L;  // one tile holds the whole lattice: periodic bonds stay inside
    const uint32_t t = lat.state[0] + t_off;
    for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) lab[i] = (int32_t)i;
    __syncthreads();
    for (uint32_t p = threadIdx.x; 2u * p < nt; p += blockDim.x) {
        const uint32_t i0 = 2u * p, ly = i0 / TW, lx0 = i0 % TW;
        const uint32_t m0 = (y0 + ly) * L + x0 + lx0;
        const u32x4 r = philox(m0 >> 1, t, 0u, kTagSwBond << 28, lat.k0, lat.k1);
        const uint32_t u[4] = {r.v0, r.v1, r.v2, r.v3};
#pragma unroll
        for (uint32_t k = 0; k < 2u; ++k) {
            const uint32_t lx = lx0 + k, m = m0 + k;
            const int8_t sm = lat.spins[m];
            if (lx + 1u < TW || wrap) {
                const uint32_t nb = (y0 + ly) * L + ((x0 + lx + 1u) & mask);
                if (sm == lat.spins[nb] && u[2u * k] < lat.bond_thr)
                    lds_unite(lab, (int32_t)(ly * TW + lx), (int32_t)(ly * TW + ((lx + 1u) & (TW - 1u))));
            }
            if (ly + 1u < TW || wrap) {
                const uint32_t nb = ((y0 + ly + 1u) & mask) * L + x0 + lx;
                if (sm == lat.spins[nb] && u[2u * k + 1u] < lat.bond_thr)
                    lds_unite(lab, (int32_t)(ly * TW + lx), (int32_t)(((ly + 1u) & (TW - 1u)) * TW + lx));
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) {
        const uint32_t rt = (uint32_t)lds_root(lab, (int32_t)i);
        lat.labels[(y0 + i / TW) * L + x0 + i % TW] = (int32_t)((y0 + rt / TW) * L + x0 + rt % TW);
    }
}

// Bonds leaving a tile: thread b < L^2/32 takes the right bond of a site on a
// tile's right edge, the rest the down bond of a site on a bottom edge.
__global__ __launch_bounds__(256) void k_sw_boundary(const HbmLattice *__restrict__ lats, uint32_t t_off)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L, mask = L - 1u;
    const uint32_t half = L * (L / kSwTile);
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 2u * half) return;
    const bool down = b >= half;
    const uint32_t c = down ? b - half : b;
    const uint32_t edge = (c / L) * kSwTile + kSwTile - 1u, along = c % L;
    const uint32_t x = down ? along : edge, y = down ? edge : along;
    const uint32_t m = y * L + x;
    const uint32_t nb = down ? ((y + 1u) & mask) * L + x : y * L + ((x + 1u) & mask);
    const uint32_t t = lat.state[0] + t_off;
    const u32x4 r = philox(m >> 1, t, 0u, kTagSwBond << 28, lat.k0, lat.k1);
    const uint32_t u = pick4(r, 2u * (m & 1u) + (down ? 1u : 0u));
    if (lat.spins[m] == lat.spins[nb] && u < lat.bond_thr) sw_unite(lat.labels, (int32_t)m, (int32_t)nb);
}

__global__ __launch_bounds__(256) void k_sw_flip(const HbmLattice *__restrict__ lats, uint32_t t_off)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t n = (uint32_t)lat.L * (uint32_t)lat.L;
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const uint32_t t = lat.state[0] + t_off;
    const uint32_t root = (uint32_t)sw_root(lat.labels, (int32_t)m);
    const u32x4 r = philox(root >> 7, t, 0u, kTagSwFlip << 28, lat.k0, lat.k1);
    if ((pick4(r, (root >> 5) & 3u) >> (root & 31u)) & 1u) lat.spins[m] = (int8_t)(-lat.spins[m]);
}

// ---------------------------------------------------------------- sweep
// One thread per word of 32 same-colour sites.  For L >= 64 a word is the
// sites of one colour in a 64-byte run of one row.
__device__ __forceinline__ void load64(const int8_t *p, uint32_t d[16])
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

__device__ __forceinline__ void store64(int8_t *p, const uint32_t d[16])
{
    uint4 *v = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = make_uint4(d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]);
}

__device__ __forceinline__ uint32_t sbit(int8_t v) { return ((uint32_t)(uint8_t)v) >> 7; }

__global__ __launch_bounds__(256) void k_int8_half_sweep(const HbmLattice *__restrict__ lats,
                                                         uint32_t t_off, uint32_t colour)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L;
    const uint32_t t = lat.state[0] + t_off;
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const Accept acc = lat.acc;
    int8_t *spins = lat.spins;
    if (L >= 64) {
        // One thread per (strip of R rows, 64-column chunk): it slides a window
        // of row chunks down its strip, so each row is read (R + 2) / R times
        // instead of three.  Only the other colour's bytes of the rows
        // above and below are used, and no thread writes those in this half.
        const uint32_t wpr = L >> 6, sh = (uint32_t)__builtin_ctz(wpr);
        const uint32_t R = half_sweep_rows(L);
        if (w >= (L / R) * wpr) return;
        const uint32_t q = w & (wpr - 1u), y0 = (w >> sh) * R;
        uint32_t U[16], C[16], D[16], E[16];
        load64(spins + (size_t)((y0 - 1u) & (L - 1u)) * L + 64u * q, U);
        load64(spins + (size_t)y0 * L + 64u * q, C);
        load64(spins + (size_t)((y0 + 1u) & (L - 1u)) * L + 64u * q, D);
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t y = y0 + r;
            // the row after next is fetched while this row is decided
            if (r + 1u < R) load64(spins + (size_t)((y + 2u) & (L - 1u)) * L + 64u * q, E);
            const uint32_t p = (y + colour) & 1u;
            const int8_t *rowbase = spins + (size_t)y * L;
            const uint32_t s = gather_parity(C, p);
            const uint32_t o = gather_parity(C, p ^ 1u);
            const uint32_t nu = gather_parity(U, p);
            const uint32_t nd = gather_parity(D, p);
            // the wrap-around site's byte: the last of the chunk before on
            // p = 0 rows, the first of the chunk after on p = 1 rows
            const uint32_t e = sbit(rowbase[(64u * q + 65u * p - 1u) & (L - 1u)]);
            const uint32_t nr = p == 0u ? o : (o >> 1) | (e << 31);
            const uint32_t nl = p == 0u ? (o << 1) | e : o;
            const uint32_t flip = metropolis_word(s, nu, nd, nl, nr, 0xFFFFFFFFu, acc, lat.k0, lat.k1,
                                                  y * wpr + q, t, colour);
            if (flip) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t f0 = (flip >> (2 * i)) & 1u, f1 = (flip >> (2 * i + 1)) & 1u;
                    C[i] ^= ((f0 * 0xFEu) << (8u * p)) | ((f1 * 0xFEu) << (8u * p + 16u));
                }
                store64(spins + (size_t)y * L + 64u * q, C);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                U[i] = C[i];
                C[i] = D[i];
                D[i] = E[i];
            }
        }
        return;
    }
    // L < 64: a word spans several rows; site-by-site.
    const uint32_t half = L >> 1;
    const uint32_t n_sites = L * half;
    const uint32_t n_words = (n_sites + 31u) >> 5;
    if (w >= n_words) return;
    uint32_t s = 0, nu = 0, nd = 0, nl = 0, nr = 0, valid = 0;
    for (uint32_t j = 0; j < 32u; ++j) {
        const uint32_t m = 32u * w + j;
        if (m >= n_sites) break;
        const uint32_t y = m / half, k = m % half;
        const uint32_t x = 2u * k + ((y + colour) & 1u);
        const uint32_t ym = (y - 1u) & (L - 1u), yp = (y + 1u) & (L - 1u);
        const uint32_t xm = (x - 1u) & (L - 1u), xp = (x + 1u) & (L - 1u);
        valid |= 1u << j;
        s |= sbit(spins[y * L + x]) << j;
        nu |= sbit(spins[ym * L + x]) << j;
        nd |= sbit(spins[yp * L + x]) << j;
        nl |= sbit(spins[y * L + xm]) << j;
        nr |= sbit(spins[y * L + xp]) << j;
    }
    const uint32_t flip = metropolis_word(s, nu, nd, nl, nr, valid, acc, lat.k0, lat.k1, w, t, colour);
    for (uint32_t j = 0; j < 32u; ++j) {
        if (!((flip >> j) & 1u)) continue;
        const uint32_t m = 32u * w + j;
        const uint32_t y = m / half, k = m % half;
        const uint32_t x = 2u * k + ((y + colour) & 1u);
        spins[y * L + x] = (int8_t)(-spins[y * L + x]);
    }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Adds the N per-thread counts v[0..N-1] of a 256-thread workgroup into
// dst[0..N-1]: wave sums, then one atomic per counter per workgroup.  Every
// thread of the workgroup calls it.
template <int N>
__device__ __forceinline__ void block_add(unsigned long long *dst, uint32_t (&v)[N])
{
    __shared__ uint32_t part[N][4];
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63u) == 0u) {
#pragma unroll
        for (int k = 0; k < N; ++k) part[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)N) {
        const uint32_t k = threadIdx.x;
        const uint32_t tot = part[k][0] + part[k][1] + part[k][2] + part[k][3];
        if (tot) atomicAdd(dst + k, (unsigned long long)tot);
    }
}

// ---------------------------------------------------------------- level 0
// Disagreement counts: dNN = #{bonds with unequal spins} over the 2L^2 nearest
// bonds, dNNN likewise over the 2L^2 diagonal bonds, dP = #{plaquettes with an
// odd number of -1 spins}, ndown = #{-1 spins}.  Then NN = 2L^2 - 2 dNN,
// NNN = 2L^2 - 2 dNNN, P = L^2 - 2 dP, M = L^2 - 2 ndown.
// Sign bits of 32 int8 spins (bit i = spin x0 + i is -1).
__device__ __forceinline__ uint32_t row_bits32(const int8_t *p)
{
    const uint4 *v = reinterpret_cast<const uint4 *>(p);
    const uint4 a = v[0], b = v[1];
    return sign4(a.x) | (sign4(a.y) << 4) | (sign4(a.z) << 8) | (sign4(a.w) << 12) | (sign4(b.x) << 16) |
           (sign4(b.y) << 20) | (sign4(b.z) << 24) | (sign4(b.w) << 28);
}

// One item per (block row by, 32-column chunk q), L >= 32: the operators of
// the sites of rows 2by and 2by + 1 in the chunk and the 16 blocked spins of
// the chunk, stored as half-word by * (L/32) + q of level 1 (bit offset
// 16 (by * L/32 + q) = by * L/2 + 16 q, as in row-major bits).
__global__ __launch_bounds__(256) void k_int8_measure0(const HbmLattice *__restrict__ lats,
                                                       uint32_t t_off)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L;
    const uint32_t t = lat.state[0] + t_off;
    const int8_t *spins = lat.spins;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c[4] = {0u, 0u, 0u, 0u};  // dNN, dNNN, dP, ndown
    if (L >= 32) {
        const uint32_t cpr = L >> 5;  // 32-site chunks per row == level-1 half-words per block row
        const uint32_t n_items = (L >> 1) * cpr;
        if (i < n_items) {
            const uint32_t by = i >> __builtin_ctz(cpr), q = i & (cpr - 1u);
            uint32_t a[3], nx[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const uint32_t y = (2u * by + (uint32_t)r) & (L - 1u);
                const int8_t *rowp = spins + (size_t)y * L;
                a[r] = row_bits32(rowp + 32u * q);
                nx[r] = sbit(rowp[(32u * q + 32u) & (L - 1u)]);
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                // right = s(x+1): the row shifted right by one with the next chunk's first bit.
                const uint32_t s = a[r], d = a[r + 1];
                const uint32_t rt = (s >> 1) | (nx[r] << 31), g = (d >> 1) | (nx[r + 1] << 31);
                c[0] += popc(s ^ rt) + popc(s ^ d);
                c[1] += popc(s ^ g) + popc(rt ^ d);
                c[2] += popc(xor3(s, rt, d) ^ g);
                c[3] += popc(s);
            }
            // coin word i >> 1 of level 0, its half i & 1 (the half-word's bit offset mod 32)
            const uint32_t coin = tie_coin_word(i >> 1, t, 0u, lat.k0, lat.k1) >> (16u * (i & 1u));
            reinterpret_cast<uint16_t *>(lat.lvl)[i] = (uint16_t)block16(a[0], a[1], coin & 0xFFFFu);
        }
    } else {
        // L < 32: one thread per level-1 word, site by site.
        const uint32_t Lb = L >> 1;
        const uint32_t nb = Lb * Lb;
        const uint32_t n_words = (nb + 31u) >> 5;
        if (i < n_words) {
            uint32_t out = 0;
            const uint32_t coin = tie_coin_word(i, t, 0u, lat.k0, lat.k1);
            for (uint32_t j = 0; j < 32u; ++j) {
                const uint32_t bm = 32u * i + j;
                if (bm >= nb) break;
                const uint32_t by = bm / Lb, bx = bm % Lb;
                uint32_t nd = 0;
                for (uint32_t v = 0; v < 2u; ++v) {
                    for (uint32_t u = 0; u < 2u; ++u) {
                        const uint32_t x = 2u * bx + u, y = 2u * by + v;
                        const uint32_t xp = (x + 1u) & (L - 1u), yp = (y + 1u) & (L - 1u);
                        const uint32_t a = sbit(spins[y * L + x]), b = sbit(spins[y * L + xp]);
                        const uint32_t cc = sbit(spins[yp * L + x]), d = sbit(spins[yp * L + xp]);
                        c[0] += (a ^ b) + (a ^ cc);
                        c[1] += (a ^ d) + (b ^ cc);
                        c[2] += a ^ b ^ cc ^ d;
                        c[3] += a;
                        nd += a;
                    }
                }
                const uint32_t bit = nd >= 3u ? 1u : nd <= 1u ? 0u : ((coin >> j) & 1u);
                out |= bit << j;
            }
            lat.lvl[i] = out;
        }
    }
    block_add<4>(lat.sums, c);
}

// ---------------------------------------------------------------- levels >= 1
// A grid-measured level l (side Ll >= kGridLevelMinL): one thread per word of
// the level for its operators, and one per word of level l + 1 for the
// blocking when l < nl (ties take the coin words of level l).
__global__ __launch_bounds__(256) void k_rm_level(const HbmLattice *__restrict__ lats, uint32_t level,
                                                  uint32_t t_off)
{
    const HbmLattice &lat = lats[blockIdx.y];
    const uint32_t L = (uint32_t)lat.L;
    const uint32_t t = lat.state[0] + t_off;
    const uint32_t Ll = L >> level;
    const uint32_t *src = lat.lvl + level_offset(L, level);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c[3] = {0u, 0u, 0u};
    if (i < rm_ops_items(Ll)) rm_ops_item(src, Ll, i, c[0], c[1], c[2]);
    if (level < (uint32_t)lat.n_levels && i < rm_block_items(Ll))
        lat.lvl[level_offset(L, level + 1u) + i] = rm_block_item(src, Ll, i, tie_coin_word(i, t, level, lat.k0, lat.k1));
    block_add<3>(lat.sums + 4u + 3u * (level - 1u), c);
}

constexpr uint32_t kLevelsThreads = 512;

// Disagreement counts of one RM level, summed over the workgroup into cnt[0..2].
__device__ void rm_ops(const uint32_t *buf, uint32_t Ll, uint32_t *cnt)
{
    uint32_t dnn = 0, dnnn = 0, dp = 0;
    const uint32_t n = rm_ops_items(Ll);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) rm_ops_item(buf, Ll, i, dnn, dnnn, dp);
    dnn = wave_sum(dnn);
    dnnn = wave_sum(dnnn);
    dp = wave_sum(dp);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(cnt + 0, dnn);
        atomicAdd(cnt + 1, dnnn);
        atomicAdd(cnt + 2, dp);
    }
}

// b = 2 majority blocking of an RM level (size Ll) into dst (size Ll/2).
__device__ void rm_block(const uint32_t *src, uint32_t Ll, uint32_t *dst, uint32_t t,
                         uint32_t level, uint32_t k0, uint32_t k1)
{
    const uint32_t n_out = rm_block_items(Ll);
    for (uint32_t i = threadIdx.x; i < n_out; i += blockDim.x)
        dst[i] = rm_block_item(src, Ll, i, tie_coin_word(i, t, level, k0, k1));
}

// One workgroup per lattice: the levels below kGridLevelMinL (the first read
// from global memory, the deeper ones in LDS), then the operator values of
// every level (grid-measured levels and level 0 from the counters, which it
// clears) and the sample's terms added to its bin record, one entry per thread.
__global__ __launch_bounds__(kLevelsThreads) void k_levels_accumulate(const HbmLattice *__restrict__ lats,
                                                                      uint32_t t_off, uint32_t q_off)
{
    const HbmLattice &lat = lats[blockIdx.x];
    const uint32_t L = (uint32_t)lat.L;
    const uint32_t nl = (uint32_t)lat.n_levels;
    const uint32_t t = lat.state[0] + t_off;
    const uint32_t l0 = n_grid_levels(L, nl) + 1u;  // first level of this workgroup
    __shared__ uint32_t cnt[kMaxLevels + 1][3];
    __shared__ uint32_t bufA[level_words(kWgLevelMaxL / 2u)];
    __shared__ uint32_t bufB[level_words(kWgLevelMaxL / 4u)];
    __shared__ int64_t S[kMaxLevels + 1][3];
    __shared__ int64_t Msh;
    for (uint32_t i = threadIdx.x; i < (kMaxLevels + 1) * 3; i += blockDim.x) (&cnt[0][0])[i] = 0u;
    __syncthreads();
    if (l0 <= nl) {
        const uint32_t *cur = lat.lvl + level_offset(L, l0);
        uint32_t Ll = L >> l0;
        for (uint32_t l = l0; l <= nl; ++l) {
            rm_ops(cur, Ll, cnt[l]);
            if (l < nl) {
                uint32_t *dst = ((l - l0) & 1u) ? bufB : bufA;
                __syncthreads();
                rm_block(cur, Ll, dst, t, l, lat.k0, lat.k1);
                __syncthreads();
                cur = dst;
                Ll >>= 1;
            }
        }
    }
    __syncthreads();
    // operator values; every counter is read and cleared by one thread
    const uint32_t n_s = 3u * (nl + 1u);
    for (uint32_t i = threadIdx.x; i <= n_s; i += blockDim.x) {
        if (i == n_s) {
            Msh = (int64_t)L * L - 2 * (int64_t)lat.sums[3];
            lat.sums[3] = 0ull;
            continue;
        }
        const uint32_t l = i / 3u, a = i % 3u;
        const int64_t nlv = (int64_t)(L >> l) * (L >> l);
        int64_t d;
        if (l == 0u || l < l0) {
            unsigned long long *sp = l == 0u ? lat.sums + a : lat.sums + 4u + 3u * (l - 1u) + a;
            d = (int64_t)*sp;
            *sp = 0ull;
        } else {
            d = cnt[l][a];
        }
        S[l][a] = (a == 2u ? nlv : 2 * nlv) - 2 * d;
    }
    __syncthreads();
    const uint32_t q = lat.state[1] + q_off;
    const uint32_t bin = q / (uint32_t)lat.sweeps_per_bin;
    const uint32_t rs = record_size(nl);
    int64_t *rec = lat.records + (size_t)bin * rs;
    for (uint32_t j = threadIdx.x; j < rs; j += blockDim.x) rec[j] += record_term(j, nl, S, Msh);
}

// Advances the chain's sweep counter and sample counter after a chunk.
__global__ void k_advance(const HbmLattice *__restrict__ lats, uint32_t n_lats, uint32_t sweeps,
                          uint32_t samples)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_lats) return;
    lats[i].state[0] += sweeps;
    lats[i].state[1] += samples;
}

// ---------------------------------------------------------------- launchers
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

hipError_t launch_int8_init(const HbmLattice *d, int n_lats, int L, hipStream_t s)
{
    const uint32_t n = (uint32_t)L * (uint32_t)L;
    hipLaunchKernelGGL(k_int8_init, dim3(cdiv(n, 256u), (uint32_t)n_lats), dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_sw_sweep(const HbmLattice *d, int n_lats, int L, uint32_t t_off, hipStream_t s)
{
    const uint32_t n = (uint32_t)L * (uint32_t)L;
    const uint32_t tw = (uint32_t)L < kSwTile ? (uint32_t)L : kSwTile;
    hipLaunchKernelGGL(k_sw_local, dim3(n / (tw * tw), (uint32_t)n_lats), dim3(256), 0, s, d, t_off);
    if ((uint32_t)L > kSwTile) {
        const uint32_t nb = 2u * (uint32_t)L * ((uint32_t)L / kSwTile);
        hipLaunchKernelGGL(k_sw_boundary, dim3(cdiv(nb, 256u), (uint32_t)n_lats), dim3(256), 0, s, d, t_off);
    }
    hipLaunchKernelGGL(k_sw_flip, dim3(cdiv(n, 256u), (uint32_t)n_lats), dim3(256), 0, s, d, t_off);
    return hipGetLastError();
}

hipError_t launch_int8_half_sweep(const HbmLattice *d, int n_lats, int L, uint32_t t_off,
                                  uint32_t colour, hipStream_t s)
{
    const uint32_t n_sites = (uint32_t)L * (uint32_t)L / 2u;
    const uint32_t n_words = cdiv(n_sites, 32u);
    // L >= 64: one thread per strip of half_sweep_rows(L) rows and 64 columns
    const uint32_t n_items = L >= 64 ? n_words / half_sweep_rows((uint32_t)L) : n_words;
    hipLaunchKernelGGL(k_int8_half_sweep, dim3(cdiv(n_items, 256u), (uint32_t)n_lats), dim3(256), 0,
                       s, d, t_off, colour);
    return hipGetLastError();
}

hipError_t launch_int8_measure0(const HbmLattice *d, int n_lats, int L, uint32_t t_off,
                                hipStream_t s)
{
    const uint32_t Lb = (uint32_t)L / 2u;
    const uint32_t n_items = L >= 32 ? Lb * ((uint32_t)L / 32u) : cdiv(Lb * Lb, 32u);
    hipLaunchKernelGGL(k_int8_measure0, dim3(cdiv(n_items, 256u), (uint32_t)n_lats), dim3(256), 0,
                       s, d, t_off);
    return hipGetLastError();
}

hipError_t launch_levels_accumulate(const HbmLattice *d, int n_lats, int L, int n_levels, uint32_t t_off,
                                    uint32_t q_off, hipStream_t s)
{
    const uint32_t g = n_grid_levels((uint32_t)L, (uint32_t)n_levels);
    for (uint32_t l = 1; l <= g; ++l) {
        const uint32_t Ll = (uint32_t)L >> l;
        hipLaunchKernelGGL(k_rm_level, dim3(cdiv(Ll * (Ll / 32u), 256u), (uint32_t)n_lats), dim3(256), 0, s, d,
                           l, t_off);
    }
    hipLaunchKernelGGL(k_levels_accumulate, dim3((uint32_t)n_lats), dim3(kLevelsThreads), 0, s, d, t_off,
                       q_off);
    return hipGetLastError();
}

hipError_t launch_advance(const HbmLattice *d, int n_lats, uint32_t sweeps, uint32_t samples,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_advance, dim3(cdiv((uint32_t)n_lats, 64u)), dim3(64), 0, s, d,
                       (uint32_t)n_lats, sweeps, samples);
    return hipGetLastError();
}

}  // namespace mcrg
