// Host FP64 part of the MCRG path: linearised RG matrix, its leading eigenvalue,
// the thermal exponent with a delete-one-bin jackknife, and the critical-point
// matching step.  Everything starts from the exact int64 per-bin records; the
// connected correlators are formed exactly in 128-bit integers before the
// conversion to double.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hbm_kernels.hpp"
#include "runtime.hpp"

namespace mcrg {
namespace {

using i128 = __int128;

// Record totals in 128-bit integers: one bin of int64 entries is exact by
// construction (jobs are admitted only while sweeps_per_bin (2L^2)^2 < 2^63),
// but sums over many bins or replicas can pass 2^63 (NN self sums near K_c are
// about 2 L^4 per sample: at L = 2048 and 1e6 samples, 3.5e19).
struct Sums {
    const i128 *r;
    uint32_t nl;
    i128 S(uint32_t l, uint32_t a) const { return r[off_S(l, a)]; }
    i128 self(uint32_t l, uint32_t a, uint32_t b) const { return r[off_self(nl, l, a, b)]; }
    i128 cross(uint32_t l, uint32_t a, uint32_t b) const { return r[off_cross(nl, l, a, b)]; }
    i128 nn0(uint32_t l) const { return r[off_nn0(nl, l)]; }
    i128 n() const { return r[off_M(nl) + 3u]; }
};

// Sum of n_bins records of record_size(nl) entries each.
std::vector<i128> total_of(const int64_t *records, int64_t n_bins, uint32_t nl)
{
    const uint32_t rs = record_size(nl);
    std::vector<i128> t(rs, 0);
    for (int64_t b = 0; b < n_bins; ++b)
        for (uint32_t k = 0; k < rs; ++k) t[k] += records[(size_t)b * rs + k];
    return t;
}

// T = A^-1 B for A, B of size n <= 3 (row-major), by the adjugate of A.
bool solve_small(int n, const double *A, const double *B, double *T)
{
    double inv[9];
    if (n == 1) {
        if (A[0] == 0.0) return false;
        inv[0] = 1.0 / A[0];
    } else if (n == 2) {
        const double det = A[0] * A[3] - A[1] * A[2];
        if (det == 0.0) return false;
        inv[0] = A[3] / det;
        inv[1] = -A[1] / det;
        inv[2] = -A[2] / det;
        inv[3] = A[0] / det;
    } else {
        const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8],
                     c02 = A[3] * A[7] - A[4] * A[6];
        const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
        if (det == 0.0) return false;
        inv[0] = c00 / det;
        inv[3] = c01 / det;
        inv[6] = c02 / det;
        inv[1] = (A[2] * A[7] - A[1] * A[8]) / det;
        inv[4] = (A[0] * A[8] - A[2] * A[6]) / det;
        inv[7] = (A[1] * A[6] - A[0] * A[7]) / det;
        inv[2] = (A[1] * A[5] - A[2] * A[4]) / det;
        inv[5] = (A[2] * A[3] - A[0] * A[5]) / det;
        inv[8] = (A[0] * A[4] - A[1] * A[3]) / det;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += inv[i * n + k] * B[k * n + j];
            T[i * n + j] = s;
        }
    return true;
}

double cubic(double a, double b, double c, double x) { return ((x + a) * x + b) * x + c; }

// Largest real part among the eigenvalues of a real n x n matrix, n <= 3.
double leading_eigenvalue(int n, const double *T)
{
    if (n == 1) return T[0];
    if (n == 2) {
        const double h = 0.5 * (T[0] + T[3]);
        const double d = h * h - (T[0] * T[3] - T[1] * T[2]);
        return d >= 0.0 ? h + std::sqrt(d) : h;
    }
    // characteristic polynomial x^3 + a x^2 + b x + c
    const double tr = T[0] + T[4] + T[8];
    const double m2 = T[0] * T[4] - T[1] * T[3] + T[0] * T[8] - T[2] * T[6] + T[4] * T[8] - T[5] * T[7];
    const double det = T[0] * (T[4] * T[8] - T[5] * T[7]) - T[1] * (T[3] * T[8] - T[5] * T[6]) +
                       T[2] * (T[3] * T[7] - T[4] * T[6]);
    const double a = -tr, b = m2, c = -det;
    // every root lies within R of the origin (Cauchy bound)
    const double R = 1.0 + std::fmax(std::fabs(a), std::fmax(std::fabs(b), std::fabs(c)));
    // bracket the largest real root: above the larger critical point when the
    // cubic dips below zero there, otherwise it is the single real root
    double lo = -R, hi = R;
    const double disc = a * a - 3.0 * b;
    if (disc > 0.0) {
        const double xp = (-a + std::sqrt(disc)) / 3.0;
        if (cubic(a, b, c, xp) <= 0.0) lo = xp;
        else hi = (-a - std::sqrt(disc)) / 3.0;
    }
    for (int it = 0; it < 200 && hi - lo > 1e-15 * (1.0 + std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (cubic(a, b, c, mid) > 0.0) hi = mid;
        else lo = mid;
    }
    double r = 0.5 * (lo + hi);
    for (int it = 0; it < 2; ++it) {
        const double f = cubic(a, b, c, r), df = (3.0 * r + 2.0 * a) * r + b;
        if (df != 0.0) r -= f / df;
    }
    // deflate: x^3 + a x^2 + b x + c = (x - r)(x^2 + p x + q)
    const double p = a + r, q = b + r * p;
    const double dq = 0.25 * p * p - q;
    const double other = dq >= 0.0 ? -0.5 * p + std::sqrt(dq) : -0.5 * p;
    return r > other ? r : other;
}

// n sum(xy) - sum(x) sum(y) as a double.  While every total is below 2^63 in
// magnitude (at L = 16384 that is 30 samples per bin times 2^31 bins) the
// products and their difference are exact 128-bit integers; beyond, they are
// formed in long double, whose 64-bit mantissa keeps the difference's relative
// error near 2^-64 of the terms, far below any statistical error.
double connected(i128 n, i128 sxy, i128 sx, i128 sy)
{
    const i128 lim = (i128)1 << 63;
    auto small = [&](i128 v) { return v < lim && v > -lim; };
    if (small(n) && small(sxy) && small(sx) && small(sy)) return (double)(n * sxy - sx * sy);
    return (double)((long double)n * (long double)sxy - (long double)sx * (long double)sy);
}

bool rg_matrix(const i128 *rec, uint32_t nl, uint32_t level, int n_ops, double *T)
{
    const Sums s{rec, nl};
    const i128 n = s.n();
    if (n < 2) return false;
    double A[9], B[9];
    for (int a = 0; a < n_ops; ++a)
        for (int b = 0; b < n_ops; ++b) {
            const i128 hi_a = s.S(level + 1, (uint32_t)a), hi_b = s.S(level + 1, (uint32_t)b);
            const i128 lo_b = s.S(level, (uint32_t)b);
            A[a * n_ops + b] = connected(n, s.self(level + 1, (uint32_t)a, (uint32_t)b), hi_a, hi_b);
            B[a * n_ops + b] = connected(n, s.cross(level + 1, (uint32_t)a, (uint32_t)b), hi_a, lo_b);
        }
    return solve_small(n_ops, A, B, T);
}

bool exponent(const i128 *rec, uint32_t nl, uint32_t level, int n_ops, double &lambda)
{
    double T[9];
    if (!rg_matrix(rec, nl, level, n_ops, T)) return false;
    lambda = leading_eigenvalue(n_ops, T);
    return lambda > 0.0;
}

// Critical coupling estimate of approx_critical_point (src/mcrg.cpp:284-299) at
// level n from the totals of an L chain (nl_l levels) and an L/2 chain.  The
// reference sums double-counted NN (2 x single count) in the numerator and in
// both covariances, so the step is formed from those values.
bool kc_step(const i128 *rl, uint32_t nll, const i128 *rs, uint32_t nls, uint32_t n, double &dK)
{
    const Sums L{rl, nll}, S{rs, nls};
    const i128 nL = L.n(), nS = S.n();
    if (nL < 2 || nS < 2) return false;
    const i128 sl = 2 * L.S(0, 0), slb = 2 * L.S(n + 1, 0), slb_sl = 4 * L.nn0(n + 1);
    const i128 ss = 2 * S.S(0, 0), ssb = 2 * S.S(n, 0);
    const i128 ssb_ss = 4 * (n == 0 ? S.self(0, 0, 0) : S.nn0(n));
    const double mean_lb = (double)slb / (double)nL, mean_sb = (double)ssb / (double)nS;
    const double cov_l = connected(nL, slb_sl, slb, sl) / ((double)nL * (double)nL);
    const double cov_s = connected(nS, ssb_ss, ssb, ss) / ((double)nS * (double)nS);
    if (cov_l - cov_s == 0.0) return false;
    dK = (mean_lb - mean_sb) / (cov_l - cov_s);
    return true;
}

}  // namespace
}  // namespace mcrg

using namespace mcrg;

extern "C" {

mcrg_status mcrg_rg_matrix(const int64_t *record, int32_t n_levels, int32_t level, int32_t n_ops, double *T)
{
    MCRG_REQUIRE(record != nullptr && T != nullptr, "null argument");
    MCRG_REQUIRE(n_levels >= 1 && n_levels <= (int32_t)kMaxLevels, "n_levels out of range");
    MCRG_REQUIRE(level >= 0 && level < n_levels, "level must be in [0, n_levels)");
    MCRG_REQUIRE(n_ops >= 1 && n_ops <= 3, "n_ops must be 1, 2 or 3");
    const std::vector<i128> tot = total_of(record, 1, (uint32_t)n_levels);
    if (!rg_matrix(tot.data(), (uint32_t)n_levels, (uint32_t)level, n_ops, T)) {
        set_error("RG matrix undefined (fewer than 2 samples or singular correlations)");
        return MCRG_ERR_INVALID_ARG;
    }
    return MCRG_OK;
}

double mcrg_leading_eigenvalue(int32_t n, const double *T)
{
    if (n < 1 || n > 3 || T == nullptr) return NAN;
    return leading_eigenvalue(n, T);
}

mcrg_status mcrg_rg_exponent(const int64_t *records, int32_t n_bins, int32_t n_levels, int32_t level,
                             int32_t n_ops, mcrg_rg_result *out)
{
    MCRG_REQUIRE(records != nullptr && out != nullptr, "null argument");
    MCRG_REQUIRE(n_bins >= 1, "n_bins must be >= 1");
    MCRG_REQUIRE(n_levels >= 1 && n_levels <= (int32_t)kMaxLevels, "n_levels out of range");
    MCRG_REQUIRE(level >= 0 && level < n_levels, "level must be in [0, n_levels)");
    MCRG_REQUIRE(n_ops >= 1 && n_ops <= 3, "n_ops must be 1, 2 or 3");
    const uint32_t nl = (uint32_t)n_levels, rs = record_size(nl);
    const std::vector<i128> tot = total_of(records, n_bins, nl);
    std::vector<i128> loo(rs);
    double lam = 0.0;
    if (!exponent(tot.data(), nl, (uint32_t)level, n_ops, lam)) {
        set_error("RG exponent undefined (too few samples, singular correlations or lambda <= 0)");
        return MCRG_ERR_INVALID_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    out->lambda = lam;
    out->yt = std::log(lam) / std::log(2.0);
    out->nu = 1.0 / out->yt;
    if (n_bins >= 2) {
        std::vector<double> yt((size_t)n_bins), nu((size_t)n_bins);
        double my = 0.0, mn = 0.0;
        int used = 0;
        for (int32_t b = 0; b < n_bins; ++b) {
            for (uint32_t k = 0; k < rs; ++k) loo[k] = tot[k] - records[(size_t)b * rs + k];
            double l = 0.0;
            if (!exponent(loo.data(), nl, (uint32_t)level, n_ops, l)) continue;
            yt[(size_t)used] = std::log(l) / std::log(2.0);
            nu[(size_t)used] = 1.0 / yt[(size_t)used];
            my += yt[(size_t)used];
            mn += nu[(size_t)used];
            ++used;
        }
        if (used >= 2) {
            my /= used;
            mn /= used;
            double sy = 0.0, sn = 0.0;
            for (int i = 0; i < used; ++i) {
                sy += (yt[(size_t)i] - my) * (yt[(size_t)i] - my);
                sn += (nu[(size_t)i] - mn) * (nu[(size_t)i] - mn);
            }
            out->yt_err = std::sqrt((double)(used - 1) / used * sy);
            out->nu_err = std::sqrt((double)(used - 1) / used * sn);
        }
        out->n_jackknife = used;
    }
    return MCRG_OK;
}

mcrg_status mcrg_tau_int(const int64_t *records, int32_t n_bins, int32_t n_levels, int32_t level, int32_t op,
                         mcrg_tau_result *out)
{
    MCRG_REQUIRE(records != nullptr && out != nullptr, "null argument");
    MCRG_REQUIRE(n_levels >= 0 && n_levels <= (int32_t)kMaxLevels, "n_levels out of range");
    MCRG_REQUIRE(op >= 0 && op <= 3, "op must be 0 (NN), 1 (NNN), 2 (P) or 3 (|M|)");
    MCRG_REQUIRE(op == 3 ? level == 0 : (level >= 0 && level <= n_levels), "level out of range");
    const uint32_t nl = (uint32_t)n_levels, rs = record_size(nl);
    const uint32_t i_sum = op == 3 ? off_M(nl) + 1u : off_S((uint32_t)level, (uint32_t)op);
    const uint32_t i_sq = op == 3 ? off_M(nl) + 2u : off_self(nl, (uint32_t)level, (uint32_t)op, (uint32_t)op);
    const uint32_t i_n = off_M(nl) + 3u;
    // full bins: the sample count of the first bin
    const int64_t c = n_bins >= 1 ? records[i_n] : 0;
    int32_t nb = 0;
    while (nb < n_bins && records[(size_t)nb * rs + i_n] == c) ++nb;
    MCRG_REQUIRE(c >= 1 && nb >= 2, "need at least two bins of equal, nonzero sample count");
    i128 s = 0, q = 0;
    std::vector<double> m((size_t)nb);
    for (int32_t b = 0; b < nb; ++b) {
        const int64_t *rb = records + (size_t)b * rs;
        s += rb[i_sum];
        q += rb[i_sq];
        m[(size_t)b] = (double)rb[i_sum] / (double)c;
    }
    const i128 n = (i128)c * nb;
    // per-sample variance from exact sums: (n q - s^2) / n^2
    const double var = connected(n, q, s, s) / ((double)n * (double)n);
    MCRG_REQUIRE(var > 0.0, "the observable does not fluctuate");
    auto tau_of = [&](const std::vector<double> &mb, double B) {
        double mean = 0.0;
        for (double x : mb) mean += x;
        mean /= (double)mb.size();
        double v = 0.0;
        for (double x : mb) v += (x - mean) * (x - mean);
        v /= (double)(mb.size() - 1);
        return 0.5 * B * v / var;
    };
    std::memset(out, 0, sizeof(*out));
    out->mean = (double)s / (double)n;
    out->var = var;
    out->tau_fine = tau_of(m, (double)c);
    // merge neighbouring bins while 32 or more remain
    double B = (double)c;
    while (m.size() / 2 >= 32) {
        std::vector<double> mm(m.size() / 2);
        for (size_t i = 0; i < mm.size(); ++i) mm[i] = 0.5 * (m[2 * i] + m[2 * i + 1]);
        m.swap(mm);
        B *= 2.0;
    }
    out->tau_int = tau_of(m, B);
    out->tau_err = out->tau_int * std::sqrt(2.0 / (double)(m.size() - 1));
    out->bin_samples = (int32_t)std::min<double>(B, 2147483647.0);
    out->n_bins = (int32_t)m.size();
    return MCRG_OK;
}

mcrg_status mcrg_critical_point_step_bins(const int64_t *rec_large, int32_t n_bins_large, int32_t nl_large,
                                          const int64_t *rec_small, int32_t n_bins_small, int32_t nl_small,
                                          int32_t level, double *dK)
{
    MCRG_REQUIRE(rec_large != nullptr && rec_small != nullptr && dK != nullptr, "null argument");
    MCRG_REQUIRE(n_bins_large >= 1 && n_bins_small >= 1, "n_bins must be >= 1");
    MCRG_REQUIRE(level >= 0 && level + 1 <= nl_large && level <= nl_small,
                 "level needs level+1 blockings of the large chain and level of the small one");
    MCRG_REQUIRE(nl_large <= (int32_t)kMaxLevels && nl_small <= (int32_t)kMaxLevels, "n_levels out of range");
    const std::vector<i128> tl = total_of(rec_large, n_bins_large, (uint32_t)nl_large);
    const std::vector<i128> ts = total_of(rec_small, n_bins_small, (uint32_t)nl_small);
    if (!kc_step(tl.data(), (uint32_t)nl_large, ts.data(), (uint32_t)nl_small, (uint32_t)level, *dK)) {
        set_error("critical-point step undefined (too few samples or equal covariances)");
        return MCRG_ERR_INVALID_ARG;
    }
    return MCRG_OK;
}

mcrg_status mcrg_critical_point_step(const int64_t *rec_large, int32_t nl_large, const int64_t *rec_small,
                                     int32_t nl_small, int32_t level, double *dK)
{
    return mcrg_critical_point_step_bins(rec_large, 1, nl_large, rec_small, 1, nl_small, level, dK);
}

int64_t mcrg_max_sweeps_per_bin(int32_t L)
{
    if (L < 2 || L > 16384) return 0;
    // the largest per-sample entry is a level-0 self product, at most (2L^2)^2
    const i128 L2 = (i128)L * L, worst = 4 * L2 * L2;
    const i128 lim = ((i128)1 << 63) - 1;
    const i128 m = lim / worst;
    return m > (i128)0x7FFFFFFF ? (int64_t)0x7FFFFFFF : (int64_t)m;
}

}  // extern "C"
