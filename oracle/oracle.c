/*
 * oracle.c -- CPU restatement of the MCRG hot path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
 * checker; never part of the product.  Conventions are documented in oracle.h.
 *
 * Written as the obvious per-site loops (no bit tricks, no laziness) so that the
 * GPU kernels' packed and lazily evaluated forms are checked against a plain
 * statement of the same rule.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11, section 4.  Replaces the reference's process-global
 * mt19937_64 (src/definitions.cpp:3-7), which is seeded from random_device and so
 * cannot be reproduced. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_seeded(uint64_t seed, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3,
                          uint32_t out[4])
{
    uint32_t ctr[4] = {x0, x1, x2, x3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* Metropolis acceptance thresholds.  The reference's Wolff sampler uses
 * P = 1 - exp(2K) (src/ising.cpp:9); a single-spin Metropolis update at the same
 * coupling accepts a move that raises sum_bonds(-s s) by 2e with probability
 * exp(-2 J e), e = s*h in {2, 4}. */
void orc_thresholds(double J, uint64_t thr[2])
{
    for (int q = 0; q < 2; ++q) {
        double p = exp(-4.0 * J * (double)(q + 1));
        double v = floor(p * 4294967296.0);
        if (v < 0.0) v = 0.0;
        if (v > 4294967296.0) v = 4294967296.0;
        thr[q] = (uint64_t)v;
    }
}

static inline int wrap(int v, int L) { return v & (L - 1); }

/* Random +-1 start, the role of Lattice::initialize_random_spins
 * (src/lattice.cpp:33-41, rand_spin() at src/definitions.hpp:34-36). */
void orc_init_spins(int L, uint64_t seed, int8_t *s)
{
    uint32_t out[4];
    for (int m = 0; m < L * L; ++m) {
        philox_seeded(seed, (uint32_t)(m >> 7), 0u, 0u, (uint32_t)ORC_TAG_INIT << 28, out);
        uint32_t bit = (out[(m >> 5) & 3] >> (m & 31)) & 1u;
        s[m] = (int8_t)(1 - 2 * (int)bit);
    }
}

/* One half-sweep over colour `color`.  This replaces one Wolff cluster
 * (IsingModel::sample_new_configuration, src/ising.cpp:87-149); neighbours follow
 * Lattice::nearest_neighbors (src/lattice.cpp:124-136).
 *
 * Random draw of site i of word w: the 4 top bits of U_i are bit i of the slices
 * out[0..3] of Philox(w, t, 0, tag) (out[0] gives bit 31).  When they equal the
 * threshold's 4 top bits the comparison is still open; the open sites of the
 * word, ranked by increasing i (rank rho = 0, 1, ...), take their 28 low bits
 * from word rho & 3 of Philox(w, t, 1 + (rho >> 2), tag) (its top 28 bits).
 * U_i is uniform on [0, 2^32) and independent across sites. */
void orc_half_sweep(int L, int8_t *s, uint64_t seed, const uint64_t thr[2], uint32_t t, int color)
{
    const int half = L / 2;
    const int n_sites = L * L / 2;
    const int n_words = (n_sites + 31) / 32;
    const uint32_t tag = ((uint32_t)ORC_TAG_SWEEP << 28) | ((uint32_t)color << 24);
    for (int w = 0; w < n_words; ++w) {
        uint32_t first[4], refine[4];
        philox_seeded(seed, (uint32_t)w, t, 0u, tag, first);
        int rank = 0;
        for (int i = 0; i < 32; ++i) {
            int m = 32 * w + i;
            if (m >= n_sites) break;
            int y = m / half;
            int k = m % half;
            int x = 2 * k + ((y + color) & 1);
            int si = s[y * L + x];
            int h = s[y * L + wrap(x + 1, L)] + s[y * L + wrap(x - 1, L)] +
                    s[wrap(y + 1, L) * L + x] + s[wrap(y - 1, L) * L + x];
            int e = si * h;
            int flip;
            if (e <= 0) {
                flip = 1;
            } else {
                uint64_t T = thr[e / 2 - 1];
                if (T >= 4294967296ull) {
                    flip = 1;
                } else {
                    uint32_t top = 0;
                    for (int kk = 0; kk < 4; ++kk) top |= ((first[kk] >> i) & 1u) << (3 - kk);
                    uint32_t ttop = (uint32_t)(T >> 28);
                    if (top != ttop) {
                        flip = top < ttop;
                    } else {
                        if ((rank & 3) == 0)
                            philox_seeded(seed, (uint32_t)w, t, 1u + (uint32_t)(rank >> 2), tag, refine);
                        uint32_t U = (top << 28) | (refine[rank & 3] >> 4);
                        flip = (uint64_t)U < T;
                        ++rank;
                    }
                }
            }
            if (flip) s[y * L + x] = (int8_t)(-si);
        }
    }
}

void orc_sweep(int L, int8_t *s, uint64_t seed, const uint64_t thr[2], uint32_t t)
{
    orc_half_sweep(L, s, seed, thr, t, 0);
    orc_half_sweep(L, s, seed, thr, t, 1);
}

/* Even operators, single count.  The reference's Lattice::calc_interactions
 * (src/lattice.cpp:102-120) sums each site against its 4 nearest and 4 diagonal
 * neighbours, which counts every bond twice; orc_ops_refcount gives that form. */
void orc_ops(int L, const int8_t *s, int64_t out[3])
{
    int64_t nn = 0, nnn = 0, p = 0;
    for (int y = 0; y < L; ++y) {
        for (int x = 0; x < L; ++x) {
            int a = s[y * L + x];
            int b = s[y * L + wrap(x + 1, L)];
            int c = s[wrap(y + 1, L) * L + x];
            int d = s[wrap(y + 1, L) * L + wrap(x + 1, L)];
            nn += a * b + a * c;
            nnn += a * d + b * c;
            p += a * b * c * d;
        }
    }
    out[0] = nn;
    out[1] = nnn;
    out[2] = p;
}

/* Direct restatement of Lattice::calc_interactions (src/lattice.cpp:102-120) with
 * the neighbour tables of src/lattice.cpp:124-151. */
void orc_ops_refcount(int L, const int8_t *s, double out[2])
{
    double s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < L; ++i) {
        for (int j = 0; j < L; ++j) {
            int nn[4][2] = {{wrap(i + 1, L), j}, {wrap(i - 1, L), j},
                            {i, wrap(j + 1, L)}, {i, wrap(j - 1, L)}};
            int nnn[4][2] = {{wrap(i + 1, L), wrap(j + 1, L)}, {wrap(i - 1, L), wrap(j + 1, L)},
                             {wrap(i + 1, L), wrap(j - 1, L)}, {wrap(i - 1, L), wrap(j - 1, L)}};
            int sij = s[j * L + i];
            for (int k = 0; k < 4; ++k) {
                s0 += sij * s[nn[k][1] * L + nn[k][0]];
                s1 += sij * s[nnn[k][1] * L + nnn[k][0]];
            }
        }
    }
    out[0] = s0;
    out[1] = s1;
}

int64_t orc_magnetization(int L, const int8_t *s)
{
    int64_t m = 0;
    for (int i = 0; i < L * L; ++i) m += s[i];
    return m;
}

/* b = 2 majority rule, MonteCarloRenormalizationGroup::block_spin_transformation
 * (src/mcrg.cpp:314-348): ties (block sum 0) take a random spin, here a Philox coin. */
void orc_block(int L, const int8_t *s, int8_t *sb, uint64_t seed, uint32_t t, int level)
{
    const int Lb = L / 2;
    for (int by = 0; by < Lb; ++by) {
        for (int bx = 0; bx < Lb; ++bx) {
            int tot = s[(2 * by) * L + 2 * bx] + s[(2 * by) * L + 2 * bx + 1] +
                      s[(2 * by + 1) * L + 2 * bx] + s[(2 * by + 1) * L + 2 * bx + 1];
            int v;
            if (tot > 0) {
                v = 1;
            } else if (tot < 0) {
                v = -1;
            } else {
                uint32_t out[4];
                int bm = by * Lb + bx;
                philox_seeded(seed, (uint32_t)(bm >> 7), t, 0u,
                              ((uint32_t)ORC_TAG_TIE << 28) | ((uint32_t)level << 16), out);
                uint32_t bit = (out[(bm >> 5) & 3] >> (bm & 31)) & 1u;
                v = 1 - 2 * (int)bit;
            }
            sb[by * Lb + bx] = (int8_t)v;
        }
    }
}

int orc_record_size(int nl) { return ORC_RECORD_SIZE(nl); }

/* One measured sample: the body of the sampling loop of
 * MonteCarloRenormalizationGroup::calc_critical_exponent (src/mcrg.cpp:72-98).
 * Row n of the reference's S holds level-n operators; Sb_S = S(n) S(n-1)^T and
 * Sb_Sb = S(n) S(n)^T are flattened column-major (src/definitions.cpp:9-19). */
void orc_measure(int L, const int8_t *s, int nl, uint64_t seed, uint32_t t, int64_t *rec)
{
    int64_t S[32][3];
    int8_t *cur = (int8_t *)malloc((size_t)L * L);
    int8_t *nxt = (int8_t *)malloc((size_t)L * L);
    memcpy(cur, s, (size_t)L * L);
    int Ll = L;
    for (int l = 0; l <= nl; ++l) {
        orc_ops(Ll, cur, S[l]);
        if (l < nl) {
            orc_block(Ll, cur, nxt, seed, t, l);
            int8_t *tmp = cur;
            cur = nxt;
            nxt = tmp;
            Ll /= 2;
        }
    }
    free(cur);
    free(nxt);
    for (int l = 0; l <= nl; ++l) {
        for (int a = 0; a < 3; ++a) {
            rec[ORC_OFF_S(nl, l, a)] += S[l][a];
            for (int b = 0; b < 3; ++b) {
                rec[ORC_OFF_SELF(nl, l, a, b)] += S[l][a] * S[l][b];
                if (l >= 1) rec[ORC_OFF_CROSS(nl, l, a, b)] += S[l][a] * S[l - 1][b];
            }
        }
        /* locator input: blocked NN against level-0 NN (src/mcrg.cpp:262-263) */
        if (l >= 1) rec[ORC_OFF_NN0(nl, l)] += S[l][0] * S[0][0];
    }
    int64_t M = orc_magnetization(L, s);
    rec[ORC_OFF_M(nl) + 0] += M;
    rec[ORC_OFF_M(nl) + 1] += M < 0 ? -M : M;
    rec[ORC_OFF_M(nl) + 2] += M * M;
    rec[ORC_OFF_N(nl)] += 1;
}

/* Swendsen-Wang bond threshold floor((1 - exp(-2J)) 2^32): an aligned pair is
 * bonded with probability 1 - exp(-2J), the cluster construction of the
 * reference's Wolff step (P_ = 1 - exp(2K), src/ising.cpp:9,134) applied to
 * every bond at once. */
uint32_t orc_sw_threshold(double J)
{
    double v = floor(-expm1(-2.0 * J) * 4294967296.0);
    if (v < 0.0) v = 0.0;
    if (v > 4294967295.0) v = 4294967295.0;
    return (uint32_t)v;
}

static int sw_find(int32_t *lab, int i)
{
    while (lab[i] != i) i = lab[i];
    return i;
}

static void sw_unite(int32_t *lab, int a, int b)
{
    a = sw_find(lab, a);
    b = sw_find(lab, b);
    if (a < b) lab[b] = a;
    else if (b < a) lab[a] = b;
}

/* One Swendsen-Wang sweep at sweep index t: the cluster move of the reference's
 * Wolff sampler (src/ising.cpp:87-149: aligned neighbours join with probability
 * P, the cluster flips) applied to all clusters at once.
 *   Bonds   site m = y*L + x owns its right bond (to (x+1, y)) and its down bond
 *           (to (x, y+1)); with out = Philox(m >> 1, t, 0, TAG_SW_BOND << 28) the
 *           right bond draws U = out[2 (m & 1)], the down bond out[2 (m & 1) + 1];
 *           a bond is active iff its two spins are equal and U < bthr.
 *   Clusters the connected components of active bonds (periodic); the root of a
 *           cluster is its smallest site index r.
 *   Flip    every site of the cluster flips iff bit (r & 31) of out[(r >> 5) & 3]
 *           of Philox(r >> 7, t, 0, TAG_SW_FLIP << 28) is 1. */
void orc_sw_sweep(int L, int8_t *s, uint64_t seed, uint32_t bthr, uint32_t t)
{
    const int n = L * L;
    int32_t *lab = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    uint32_t out[4];
    for (int m = 0; m < n; ++m) lab[m] = m;
    for (int m = 0; m < n; ++m) {
        const int x = m & (L - 1), y = m / L;
        const int right = y * L + wrap(x + 1, L), down = wrap(y + 1, L) * L + x;
        philox_seeded(seed, (uint32_t)(m >> 1), t, 0u, (uint32_t)ORC_TAG_SW_BOND << 28, out);
        if (s[m] == s[right] && out[2 * (m & 1)] < bthr) sw_unite(lab, m, right);
        if (s[m] == s[down] && out[2 * (m & 1) + 1] < bthr) sw_unite(lab, m, down);
    }
    for (int m = 0; m < n; ++m) {
        const int r = sw_find(lab, m);
        philox_seeded(seed, (uint32_t)(r >> 7), t, 0u, (uint32_t)ORC_TAG_SW_FLIP << 28, out);
        if ((out[(r >> 5) & 3] >> (r & 31)) & 1u) s[m] = (int8_t)-s[m];
    }
    free(lab);
}

/* Thermalisation then measurement, as IsingModel::equilibrate (src/ising.cpp:39-41)
 * followed by the sampling loop of calc_critical_exponent (src/mcrg.cpp:72-98).
 * Sweep indices run t0 .. t0+n_therm+n_meas-1; sample q = t - t0 - n_therm lands in
 * bin q / sweeps_per_bin.  `s` must hold the starting configuration. */
void orc_run_chain_sampler(int L, double J, uint64_t seed, uint32_t t0, int n_therm, int n_meas, int nl,
                           int sweeps_per_bin, int sampler, int8_t *s, int64_t *records)
{
    uint64_t thr[2];
    orc_thresholds(J, thr);
    const uint32_t bthr = orc_sw_threshold(J);
    const int rs = ORC_RECORD_SIZE(nl);
    for (int q = 0; q < n_therm + n_meas; ++q) {
        uint32_t t = t0 + (uint32_t)q;
        if (sampler == ORC_SAMPLER_SW)
            orc_sw_sweep(L, s, seed, bthr, t);
        else
            orc_sweep(L, s, seed, thr, t);
        if (q >= n_therm) {
            int bin = (q - n_therm) / sweeps_per_bin;
            orc_measure(L, s, nl, seed, t, records + (size_t)bin * rs);
        }
    }
}

void orc_run_chain(int L, double J, uint64_t seed, uint32_t t0, int n_therm, int n_meas,
                   int nl, int sweeps_per_bin, int8_t *s, int64_t *records)
{
    orc_run_chain_sampler(L, J, seed, t0, n_therm, n_meas, nl, sweeps_per_bin, ORC_SAMPLER_METROPOLIS, s, records);
}

/* ---------------------------------------------------------------------------
 * Linearised RG matrix (src/mcrg.cpp:110-131), from integer sums.
 * dSb_dK  = <S(n+1) S(n)^T>  - <S(n+1)><S(n)>^T
 * dSb_dKb = <S(n+1) S(n+1)^T> - <S(n+1)><S(n+1)>^T
 * T = dSb_dKb^-1 dSb_dK.  Both matrices are scaled by N^2 (N = sample count),
 * which leaves T unchanged; the numerators are exact in 128-bit integers.
 * ------------------------------------------------------------------------- */
static int solve_linear(int n, double *A, double *B)
{
    /* Solves A X = B in place (B <- X), Gaussian elimination, partial pivoting. */
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabs(A[r * n + col]) > fabs(A[piv * n + col])) piv = r;
        if (A[piv * n + col] == 0.0) return -1;
        if (piv != col) {
            for (int c = 0; c < n; ++c) {
                double t = A[col * n + c]; A[col * n + c] = A[piv * n + c]; A[piv * n + c] = t;
                t = B[col * n + c]; B[col * n + c] = B[piv * n + c]; B[piv * n + c] = t;
            }
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            double f = A[r * n + col] / A[col * n + col];
            for (int c = 0; c < n; ++c) {
                A[r * n + c] -= f * A[col * n + c];
                B[r * n + c] -= f * B[col * n + c];
            }
        }
    }
    for (int r = 0; r < n; ++r) {
        double d = A[r * n + r];
        for (int c = 0; c < n; ++c) B[r * n + c] /= d;
    }
    return 0;
}

int orc_rg_matrix(const int64_t *rec, int nl, int level, int n_ops, double *T)
{
    if (level < 0 || level >= nl || n_ops < 1 || n_ops > 3) return -1;
    const __int128 n = rec[ORC_OFF_N(nl)];
    if (n <= 1) return -1;
    double A[9], B[9];
    for (int a = 0; a < n_ops; ++a) {
        for (int b = 0; b < n_ops; ++b) {
            __int128 sa = rec[ORC_OFF_S(nl, level + 1, a)];
            __int128 sb_hi = rec[ORC_OFF_S(nl, level + 1, b)];
            __int128 sb_lo = rec[ORC_OFF_S(nl, level, b)];
            __int128 self = rec[ORC_OFF_SELF(nl, level + 1, a, b)];
            __int128 cross = rec[ORC_OFF_CROSS(nl, level + 1, a, b)];
            A[a * n_ops + b] = (double)(n * self - sa * sb_hi);
            B[a * n_ops + b] = (double)(n * cross - sa * sb_lo);
        }
    }
    if (solve_linear(n_ops, A, B) != 0) return -1;
    memcpy(T, B, sizeof(double) * (size_t)(n_ops * n_ops));
    return 0;
}

/* Largest real part over the eigenvalues of an n x n real matrix, n <= 3: the
 * reference's EigenSolver + max of real parts (src/mcrg.cpp:126-130). */
double orc_leading_eigenvalue(int n, const double *T)
{
    if (n == 1) return T[0];
    if (n == 2) {
        double tr = T[0] + T[3];
        double det = T[0] * T[3] - T[1] * T[2];
        double disc = 0.25 * tr * tr - det;
        if (disc < 0.0) return 0.5 * tr;
        return 0.5 * tr + sqrt(disc);
    }
    /* characteristic polynomial x^3 + a x^2 + b x + c */
    double tr = T[0] + T[4] + T[8];
    double m2 = (T[0] * T[4] - T[1] * T[3]) + (T[0] * T[8] - T[2] * T[6]) +
                (T[4] * T[8] - T[5] * T[7]);
    double det = T[0] * (T[4] * T[8] - T[5] * T[7]) - T[1] * (T[3] * T[8] - T[5] * T[6]) +
                 T[2] * (T[3] * T[7] - T[4] * T[6]);
    double a = -tr, b = m2, c = -det;
    double p = b - a * a / 3.0;
    double q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c;
    double D = 0.25 * q * q + p * p * p / 27.0;
    double roots[3];
    int nr;
    if (D > 0.0) {
        double sq = sqrt(D);
        double u = cbrt(-0.5 * q + sq), v = cbrt(-0.5 * q - sq);
        roots[0] = u + v - a / 3.0;
        nr = 1;
    } else {
        double r = sqrt(-p / 3.0);
        double arg = (r > 0.0) ? (-0.5 * q) / (r * r * r) : 0.0;
        if (arg > 1.0) arg = 1.0;
        if (arg < -1.0) arg = -1.0;
        double phi = acos(arg);
        for (int k = 0; k < 3; ++k)
            roots[k] = 2.0 * r * cos((phi - 2.0 * M_PI * k) / 3.0) - a / 3.0;
        nr = 3;
    }
    /* Newton polish on the cubic */
    for (int k = 0; k < nr; ++k) {
        for (int it = 0; it < 3; ++it) {
            double x = roots[k];
            double f = ((x + a) * x + b) * x + c;
            double df = (3.0 * x + 2.0 * a) * x + b;
            if (df == 0.0) break;
            roots[k] = x - f / df;
        }
    }
    if (nr == 1) {
        /* complex pair: real part = (tr - real root) / 2 */
        double re = 0.5 * (tr - roots[0]);
        return roots[0] > re ? roots[0] : re;
    }
    double best = roots[0];
    for (int k = 1; k < 3; ++k) if (roots[k] > best) best = roots[k];
    return best;
}

/* y_t = ln(lambda) / ln(b) = 1/nu, nu = ln b / ln lambda (src/mcrg.cpp:131). */
double orc_yt(const int64_t *rec, int nl, int level, int n_ops)
{
    double T[9];
    if (orc_rg_matrix(rec, nl, level, n_ops, T) != 0) return NAN;
    double lam = orc_leading_eigenvalue(n_ops, T);
    return log(lam) / log(2.0);
}

/* Delete-one-bin jackknife over per-bin records.  mean = full-sample estimate,
 * err = sqrt((nb-1)/nb * sum_i (y_(i) - ybar)^2). */
int orc_jackknife_yt(const int64_t *records, int n_bins, int nl, int level, int n_ops,
                     double *mean, double *err)
{
    const int rs = ORC_RECORD_SIZE(nl);
    int64_t *tot = (int64_t *)calloc((size_t)rs, sizeof(int64_t));
    int64_t *loo = (int64_t *)calloc((size_t)rs, sizeof(int64_t));
    double *yi = (double *)calloc((size_t)n_bins, sizeof(double));
    for (int b = 0; b < n_bins; ++b)
        for (int k = 0; k < rs; ++k) tot[k] += records[(size_t)b * rs + k];
    *mean = orc_yt(tot, nl, level, n_ops);
    double ybar = 0.0;
    for (int b = 0; b < n_bins; ++b) {
        for (int k = 0; k < rs; ++k) loo[k] = tot[k] - records[(size_t)b * rs + k];
        yi[b] = orc_yt(loo, nl, level, n_ops);
        ybar += yi[b];
    }
    ybar /= n_bins;
    double ss = 0.0;
    for (int b = 0; b < n_bins; ++b) ss += (yi[b] - ybar) * (yi[b] - ybar);
    *err = sqrt((double)(n_bins - 1) / n_bins * ss);
    free(tot);
    free(loo);
    free(yi);
    return 0;
}

/* One matching step of approx_critical_point (src/mcrg.cpp:284-299) at level n:
 * the L lattice blocked n+1 times against the L/2 lattice blocked n times, NN
 * operator, with the reference's double-counted values (2 x single count). */
int orc_kc_step(const int64_t *rl, int nll, const int64_t *rs, int nls, int n, double *dK)
{
    const double NL = (double)rl[ORC_OFF_N(nll)], NS = (double)rs[ORC_OFF_N(nls)];
    const double SL = 2.0 * rl[ORC_OFF_S(nll, 0, 0)] / NL;
    const double SS = 2.0 * rs[ORC_OFF_S(nls, 0, 0)] / NS;
    const double SLb = 2.0 * rl[ORC_OFF_S(nll, n + 1, 0)] / NL;
    const double SSb = 2.0 * rs[ORC_OFF_S(nls, n, 0)] / NS;
    const double SLb_SL = 4.0 * rl[ORC_OFF_NN0(nll, n + 1)] / NL;
    const double SSb_SS = 4.0 * (n == 0 ? rs[ORC_OFF_SELF(nls, 0, 0, 0)] : rs[ORC_OFF_NN0(nls, n)]) / NS;
    const double dSL_dK = SLb_SL - SLb * SL;
    const double dSS_dK = SSb_SS - SSb * SS;
    if (dSL_dK - dSS_dK == 0.0) return -1;
    *dK = (SLb - SSb) / (dSL_dK - dSS_dK);
    return 0;
}

/* DP work split, split_samples (src/definitions.cpp:79-87): rank 0 takes the
 * remainder (which the reference lets go negative). */
int orc_split_samples(int rank, int n_processes, int n_samples)
{
    int loc = (int)ceil((double)n_samples / (double)n_processes);
    if (rank == 0) loc = n_samples - (n_processes - 1) * loc;
    return loc;
}
