/*
 * oracle.h -- CPU restatement of the MCRG hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP implementation in mcrg_amd/.  Only the
 * test suite, the smoke check and the CPU-baseline leg of bench.py load it; the
 * product never links or calls it.  Every function is written as plain loops over
 * an int8 row-major lattice so that it can be read against the reference:
 *
 *   reference (kim-jane/MCRG)          this file
 *   --------------------------------   ------------------------------------------
 *   src/lattice.cpp:33-41  init        orc_init_spins       (Philox instead of mt19937_64)
 *   src/ising.cpp:87-149   Wolff       orc_half_sweep/sweep (checkerboard Metropolis),
 *                                      orc_sw_sweep (Swendsen-Wang: all clusters at once)
 *   src/lattice.cpp:84-120 operators   orc_ops (single count), orc_ops_refcount
 *   src/mcrg.cpp:314-348   blocking    orc_block
 *   src/mcrg.cpp:72-98     accumulate  orc_measure
 *   src/mcrg.cpp:100-137   T, lambda   orc_rg_matrix, orc_leading_eigenvalue, orc_yt
 *   src/mcrg.cpp:284-299   K_c step    orc_kc_step
 *   src/definitions.cpp:9-19 flatten   record layout (column-major 3x3 blocks)
 *   src/definitions.cpp:79-87 split    orc_split_samples
 *
 * Conventions (these define the canonical stream that the GPU must reproduce
 * bit for bit):
 *
 *  Lattice   L x L, periodic, L a power of two.  Site (x, y) is stored at
 *            s[y*L + x] (this equals the reference's column-major (i, j) at
 *            j*N + i with x = i, y = j).  s = +1 or -1.  Spin bit b = (1 - s)/2.
 *  Coupling  J = |K| > 0 (ferromagnetic).  The reference writes K < 0 for the
 *            same model (P = 1 - exp(2K), src/ising.cpp:9).  Weight exp(J * sum_bonds s s).
 *  Colours   colour(x, y) = (x + y) & 1.  Sites of colour c, enumerated
 *            m = y*(L/2) + k, x = 2k + ((y + c) & 1), k in [0, L/2).
 *            Word w = m >> 5 holds 32 sites, site bit i = m & 31.
 *  Philox    Philox4x32-10 (Salmon et al., SC'11), key = (lo32(seed), hi32(seed)),
 *            counter (x0, x1, x2, x3), output out[0..3].
 *  Sweep t   half-sweep colour 0, then colour 1.  For word w of colour c let
 *            tag = TAG_SWEEP << 28 | c << 24.  Site i of the word draws a 32-bit
 *            U_i: its 4 top bits are bit i of out[0..3] of Philox(w, t, 0, tag)
 *            (out[0] -> bit 31).  With h = sum of the 4 neighbours, e = s*h:
 *            e <= 0 -> flip; e = 2 -> flip iff U < thr[0]; e = 4 -> iff U < thr[1],
 *            thr[q] = floor(exp(-4 J (q+1)) * 2^32) (2^32 means always).  When the
 *            4 top bits of U_i equal those of thr the site is "open": the open
 *            sites of the word, ranked rho = 0, 1, ... by increasing i, take the
 *            top 28 bits of word rho & 3 of Philox(w, t, 1 + (rho >> 2), tag) as
 *            the 28 low bits of U_i.
 *  Init      site m = y*L + x gets bit (out[(m>>5)&3] >> (m&31)) & 1 of
 *            Philox(x0 = m >> 7, 0, 0, TAG_INIT << 28); s = 1 - 2 bit.
 *  Blocking  b = 2 majority rule from level l (size Ll) to level l+1 at sweep t:
 *            block (bx, by) sums s(2bx+u, 2by+v), u, v in {0,1}; > 0 -> +1,
 *            < 0 -> -1, = 0 -> coin: bm = by*(Ll/2) + bx, bit =
 *            (out[(bm>>5)&3] >> (bm&31)) & 1 of Philox(bm >> 7, t, 0,
 *            TAG_TIE << 28 | l << 16); s = 1 - 2 bit.
 *  Operators (single count; the reference's values are exactly twice NN, NNN)
 *            NN  = sum_(x,y) s(x,y) [s(x+1,y) + s(x,y+1)]
 *            NNN = sum_(x,y) [s(x,y) s(x+1,y+1) + s(x+1,y) s(x,y+1)]
 *            P   = sum_(x,y) s(x,y) s(x+1,y) s(x,y+1) s(x+1,y+1)
 *  Record    per bin, int64, for levels l = 0..nl (see ORC_* offsets below):
 *            sums of S_l, of S_l S_l^T, of S_l S_(l-1)^T (l >= 1), of
 *            NN_l NN_0 (l >= 1), of M, |M|, M^2, and the sample count.
 */
#ifndef MCRG_ORACLE_H
#define MCRG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TAG_SWEEP = 1, ORC_TAG_TIE = 2, ORC_TAG_INIT = 3, ORC_TAG_SW_BOND = 4, ORC_TAG_SW_FLIP = 5 };
enum { ORC_SAMPLER_METROPOLIS = 0, ORC_SAMPLER_SW = 1 };

/* Record offsets for a chain measured on levels 0..nl (nl = number of blockings). */
#define ORC_OFF_S(nl, l, a)        ((l) * 3 + (a))
#define ORC_OFF_SELF(nl, l, a, b)  (3 * ((nl) + 1) + (l) * 9 + (b) * 3 + (a))
#define ORC_OFF_CROSS(nl, l, a, b) (12 * ((nl) + 1) + ((l) - 1) * 9 + (b) * 3 + (a))
#define ORC_OFF_NN0(nl, l)         (12 * ((nl) + 1) + 9 * (nl) + (l) - 1)
#define ORC_OFF_M(nl)              (12 * ((nl) + 1) + 10 * (nl))
#define ORC_OFF_N(nl)              (ORC_OFF_M(nl) + 3)
#define ORC_RECORD_SIZE(nl)        (ORC_OFF_N(nl) + 1)

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_thresholds(double J, uint64_t thr[2]);

void orc_init_spins(int L, uint64_t seed, int8_t *s);
void orc_half_sweep(int L, int8_t *s, uint64_t seed, const uint64_t thr[2], uint32_t t, int color);
void orc_sweep(int L, int8_t *s, uint64_t seed, const uint64_t thr[2], uint32_t t);

void orc_ops(int L, const int8_t *s, int64_t out[3]);
void orc_ops_refcount(int L, const int8_t *s, double out[2]);
int64_t orc_magnetization(int L, const int8_t *s);
void orc_block(int L, const int8_t *s, int8_t *sb, uint64_t seed, uint32_t t, int level);

int  orc_record_size(int nl);
void orc_measure(int L, const int8_t *s, int nl, uint64_t seed, uint32_t t, int64_t *rec);
void orc_run_chain(int L, double J, uint64_t seed, uint32_t t0, int n_therm, int n_meas,
                   int nl, int sweeps_per_bin, int8_t *s, int64_t *records);
/* The same with the sweep chosen by `sampler` (ORC_SAMPLER_*). */
void orc_run_chain_sampler(int L, double J, uint64_t seed, uint32_t t0, int n_therm, int n_meas, int nl,
                           int sweeps_per_bin, int sampler, int8_t *s, int64_t *records);

uint32_t orc_sw_threshold(double J);
void orc_sw_sweep(int L, int8_t *s, uint64_t seed, uint32_t bthr, uint32_t t);

int    orc_rg_matrix(const int64_t *rec, int nl, int level, int n_ops, double *T);
double orc_leading_eigenvalue(int n, const double *T);
double orc_yt(const int64_t *rec, int nl, int level, int n_ops);
int    orc_jackknife_yt(const int64_t *records, int n_bins, int nl, int level, int n_ops,
                        double *mean, double *err);
int    orc_kc_step(const int64_t *rec_large, int nl_large, const int64_t *rec_small, int nl_small,
                   int level, double *dK);
int    orc_split_samples(int rank, int n_processes, int n_samples);

#ifdef __cplusplus
}
#endif
#endif
