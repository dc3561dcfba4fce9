// Implementation of the reference-named classes over the C ABI.
#include "mcrg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <limits>
#include <mutex>
#include <sstream>

namespace mcrg_cpp {
namespace {
std::mutex g_mu;
mcrg_ctx g_ctx = nullptr;
int g_device = 0;
uint64_t g_seed_base = 1;
uint64_t g_seed_count = 0;

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

void check(mcrg_status st)
{
    if (st != MCRG_OK) throw Error(st, std::string("mcrg: ") + mcrg_last_error());
}

void set_device(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_device = device;
}

mcrg_ctx context()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) check(mcrg_ctx_create(g_device, &g_ctx));
    return g_ctx;
}

void set_seed(uint64_t base)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_seed_base = base;
    g_seed_count = 0;
}

uint64_t next_seed()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return splitmix64(g_seed_base + g_seed_count++);
}

std::string rounded_str(double num, int precision)
{
    std::stringstream ss;
    ss << std::setprecision(precision) << num;
    return ss.str();
}

bool write_iter(int i)
{
    if (i < 10) return true;
    if (i < 100) return i % 10 == 0;
    if (i < 1000) return i % 100 == 0;
    if (i < 10000) return i % 1000 == 0;
    if (i < 100000) return i % 10000 == 0;
    return i % 100000 == 0;
}

int split_samples(int rank, int n_processes, int n_samples)
{
    int loc = (int)std::ceil((double)n_samples / (double)n_processes);
    if (rank == 0) loc = n_samples - (n_processes - 1) * loc;
    return loc;
}
}  // namespace mcrg_cpp

using mcrg_cpp::check;

namespace {
int levels_of(int N) { return std::max(0, (int)std::floor(std::log((double)N) / std::log(2.0)) - 1); }

mcrg_chain new_chain(int N, double J, uint64_t seed, int n_levels, int spb, int bins)
{
    mcrg_chain_desc d{};
    d.L = N;
    d.coupling = J;
    d.seed = seed;
    d.n_levels = n_levels;
    d.layout = MCRG_LAYOUT_INT8;
    d.sweeps_per_bin = spb;
    d.max_bins = bins;
    d.sampler = MCRG_SAMPLER_METROPOLIS;
    mcrg_chain ch = nullptr;
    check(mcrg_chain_create(mcrg_cpp::context(), &d, &ch));
    return ch;
}
}  // namespace

// ------------------------------------------------------------------ Lattice
Lattice::Lattice(int N) : N_(N), n_spins_(N * N), a_(1)
{
    seed_ = mcrg_cpp::next_seed();
    chain_ = new_chain(N, 0.0, seed_, 0, 1, 1);
    sync_spins();
}

Lattice::Lattice(int a, imat spins) : N_(spins.rows()), n_spins_(spins.rows() * spins.rows()), a_(a)
{
    spins_ = spins;
    seed_ = mcrg_cpp::next_seed();
    chain_ = new_chain(N_, 0.0, seed_, 0, 1, 1);
    push_spins();
}

Lattice::~Lattice()
{
    if (chain_) mcrg_chain_destroy(chain_);
}

void Lattice::set_coupling(double J)
{
    if (J == J_) return;
    check(mcrg_chain_set_coupling(chain_, J));
    J_ = J;
}

void Lattice::sync_spins()
{
    std::vector<int8_t> s((size_t)n_spins_);
    check(mcrg_chain_get_spins(chain_, s.data(), (int64_t)s.size()));
    spins_.resize(N_, N_);
    // device (x, y) at y*N + x == column-major (i, j) at j*N + i with i = x, j = y
    for (int j = 0; j < N_; ++j)
        for (int i = 0; i < N_; ++i) spins_(i, j) = s[(size_t)j * N_ + i];
}

void Lattice::push_spins()
{
    std::vector<int8_t> s((size_t)n_spins_);
    for (int j = 0; j < N_; ++j)
        for (int i = 0; i < N_; ++i) s[(size_t)j * N_ + i] = (int8_t)spins_(i, j);
    check(mcrg_chain_set_spins(chain_, s.data(), (int64_t)s.size()));
}

void Lattice::write_spins(FILE *fptr)
{
    sync_spins();
    std::fprintf(fptr, "\n");
    for (int i = 0; i < N_; ++i) {
        std::fprintf(fptr, "# ");
        for (int j = 0; j < N_; ++j) std::fprintf(fptr, "%i, ", spins_(i, j));
        std::fprintf(fptr, "\n");
    }
}

void Lattice::display_spins()
{
    sync_spins();
    for (int i = 0; i < N_; ++i) {
        for (int j = 0; j < N_; ++j) std::printf(spins_(i, j) == 1 ? "\033[34m%2s\033[0m" : "\033[37m%2s\033[0m",
                                                 spins_(i, j) == 1 ? "O" : "X");
        std::printf("\n");
    }
    std::printf("(Lattice spacing a = %i)\n\n", a_);
}

int Lattice::choose_random_spin()
{
    // a uniform site index from the lattice's own counter-based sequence
    uint64_t z = seed_ ^ (0xA0761D6478BD642Full * ++pick_);
    z = (z ^ (z >> 32)) * 0xD6E8FEB86659FD93ull;
    z ^= z >> 32;
    return (int)(((z >> 11) * (double)n_spins_) / 9007199254740992.0);
}

double Lattice::calc_nearest_neighbor_interaction()
{
    int64_t ops[4];
    check(mcrg_chain_operators(chain_, ops));
    return 2.0 * (double)ops[0];
}

vec2D Lattice::calc_interactions()
{
    int64_t ops[4];
    check(mcrg_chain_operators(chain_, ops));
    vec2D S;
    S(0) = 2.0 * (double)ops[0];
    S(1) = 2.0 * (double)ops[1];
    return S;
}

imat Lattice::nearest_neighbors(int i, int j)
{
    imat nn(4, 2);
    const int N = N_;
    nn(0, 0) = ((i + 1) % N + N) % N;
    nn(0, 1) = j;
    nn(1, 0) = ((i - 1) % N + N) % N;
    nn(1, 1) = j;
    nn(2, 0) = i;
    nn(2, 1) = ((j + 1) % N + N) % N;
    nn(3, 0) = i;
    nn(3, 1) = ((j - 1) % N + N) % N;
    return nn;
}

imat Lattice::next_nearest_neighbors(int i, int j)
{
    imat nnn(4, 2);
    const int N = N_;
    const int ip = ((i + 1) % N + N) % N, im = ((i - 1) % N + N) % N;
    const int jp = ((j + 1) % N + N) % N, jm = ((j - 1) % N + N) % N;
    nnn(0, 0) = ip;
    nnn(0, 1) = jp;
    nnn(1, 0) = im;
    nnn(1, 1) = jp;
    nnn(2, 0) = ip;
    nnn(2, 1) = jm;
    nnn(3, 0) = im;
    nnn(3, 1) = jm;
    return nnn;
}

// ------------------------------------------------------------------ IsingModel
IsingModel::IsingModel(double K) : K_(K)
{
    if (!(K <= 0.0)) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "IsingModel: K must be <= 0 (ferromagnetic)");
}

void IsingModel::bind(Lattice &lat) { lat.set_coupling(-K_); }

void IsingModel::sample_new_configuration(std::shared_ptr<Lattice> pLattice)
{
    bind(*pLattice);
    check(mcrg_chain_sweep(pLattice->chain(), 1, 0));
}

void IsingModel::equilibrate(std::shared_ptr<Lattice> pLattice, int n_samples_eq, bool write)
{
    bind(*pLattice);
    if (!write) {
        check(mcrg_chain_sweep(pLattice->chain(), n_samples_eq, 0));
        check(mcrg_chain_sync(pLattice->chain()));
        return;
    }
    const std::string name = "equilibrate_N_" + std::to_string(pLattice->N_) + "_K_" +
                             mcrg_cpp::rounded_str(K_, 7) + ".txt";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "cannot open " + name);
    std::fprintf(f, "# Nearest neighbor coupling K = %lf\n", K_);
    std::fprintf(f, "# Temperature T = %lf\n", -1 / K_);
    std::fprintf(f, "# Number of lattice sites = %i\n", pLattice->N_ * pLattice->N_);
    std::fprintf(f, "# Lattice spacing = %i\n", pLattice->a_);
    std::fprintf(f, "# Using %i parallel processes\n", 1);
    std::fprintf(f, "# %s, %s, %s, %s, %s, %s, %s\n", "Iteration", "Avg E/spin", "Stddev E/spin", "Heat Capacity",
                 "Avg |M|/spin", "Stddev |M|/spin", "Susceptibility");
    for (int n = 1; n <= n_samples_eq; ++n) {
        check(mcrg_chain_sweep(pLattice->chain(), 1, 0));
        if (!mcrg_cpp::write_iter(n)) continue;
        // one chain: the average over ranks of src/ising.cpp:50-63 is the value itself
        const double E = calc_energy(pLattice), M = std::fabs(calc_magnetization(pLattice));
        const double E_sigma = E * E - E * E, M_sigma = M * M - M * M;
        std::fprintf(f, "%i, %10.7e, %10.7e, %10.7e, %10.7e, %10.7e, %10.7e\n", n, E, std::sqrt(E_sigma),
                     E_sigma * K_ * K_, M, std::sqrt(M_sigma), -M_sigma * K_);
    }
    pLattice->write_spins(f);
    std::fclose(f);
}

double IsingModel::calc_energy(std::shared_ptr<Lattice> pLattice)
{
    // K * (sum over sites of s times its 4 neighbours) / N^2, src/ising.cpp:158-173
    int64_t ops[4];
    check(mcrg_chain_operators(pLattice->chain(), ops));
    return K_ * 2.0 * (double)ops[0] / ((double)pLattice->N_ * pLattice->N_);
}

double IsingModel::calc_magnetization(std::shared_ptr<Lattice> pLattice)
{
    int64_t ops[4];
    check(mcrg_chain_operators(pLattice->chain(), ops));
    return (double)ops[3] / ((double)pLattice->N_ * pLattice->N_);
}

// ------------------------------------------------------------------ MCRG
MonteCarloRenormalizationGroup::MonteCarloRenormalizationGroup(int b) : b_(b)
{
    if (b != 2) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "only b = 2 majority-rule blocking is supported");
    mcrg_cpp::context();
    std::printf("\n=============================================\n");
    std::printf("==========       MONTE CARLO       ==========\n");
    std::printf("==========  RENORMALIZATION GROUP  ==========\n");
    std::printf("=============================================\n\n");
    std::printf("* Using %i parallel process(es)\n", n_processes_);
    std::printf("* Scaling factor b = %i\n", b_);
}

void MonteCarloRenormalizationGroup::set_operators(int n_ops)
{
    if (n_ops < 1 || n_ops > 3) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "n_ops must be 1, 2 or 3");
    n_ops_ = n_ops;
}
void MonteCarloRenormalizationGroup::set_replicas(int replicas)
{
    if (replicas < 1) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "replicas must be >= 1");
    n_processes_ = replicas;
}
void MonteCarloRenormalizationGroup::set_bins(int bins)
{
    if (bins < 1) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "bins must be >= 1");
    bins_ = bins;
}
void MonteCarloRenormalizationGroup::set_levels(int levels) { levels_ = levels; }
void MonteCarloRenormalizationGroup::set_output_dir(const std::string &dir) { dir_ = dir; }
void MonteCarloRenormalizationGroup::set_quiet(bool quiet) { quiet_ = quiet; }
void MonteCarloRenormalizationGroup::set_sampler(int sampler)
{
    if (sampler != MCRG_SAMPLER_METROPOLIS && sampler != MCRG_SAMPLER_SWENDSEN_WANG)
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "unknown sampler");
    sampler_ = sampler;
}

void MonteCarloRenormalizationGroup::set_layout(int layout)
{
    if (layout != -1 && layout != MCRG_LAYOUT_INT8 && layout != MCRG_LAYOUT_PACKED)
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "layout must be -1 (auto), MCRG_LAYOUT_INT8 or MCRG_LAYOUT_PACKED");
    layout_ = layout;
}

std::vector<int64_t> MonteCarloRenormalizationGroup::run_chains(int N, double K, int n_levels, int n_samples_eq,
                                                                int n_samples, int &n_bins)
{
    const int R = n_processes_;
    const int rs = mcrg_record_size(n_levels);
    std::vector<int> n_loc((size_t)R);
    int max_loc = 0;
    for (int r = 0; r < R; ++r) {
        n_loc[(size_t)r] = std::max(0, mcrg_cpp::split_samples(r, R, n_samples));
        max_loc = std::max(max_loc, n_loc[(size_t)r]);
    }
    // Bin b of the result is the sum of bin b of every replica, kept in int64:
    // each bin holds at most R * spb samples, so spb is capped at
    // mcrg_max_sweeps_per_bin(N) / R (more bins than asked for when needed).
    const int64_t spb_cap = mcrg_max_sweeps_per_bin(N) / R;
    if (spb_cap < 1)
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG,
                              "too many replicas at this N: one summed sample per replica would overflow int64");
    const int spb = (int)std::min<int64_t>(std::max(1, (max_loc + bins_ - 1) / bins_), spb_cap);
    n_bins = std::max(1, (max_loc + spb - 1) / spb);
    std::vector<int64_t> total((size_t)n_bins * rs, 0);
    const double J = -K;
    if (layout_ == MCRG_LAYOUT_PACKED && (N < 8 || N > 512 || sampler_ != MCRG_SAMPLER_METROPOLIS))
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "the bit-packed layout runs Metropolis at 8 <= N <= 512 only");
    if (layout_ == MCRG_LAYOUT_INT8) {
        // one int8 chain per replica, each on its own stream
        std::vector<mcrg_chain> chs((size_t)R, nullptr);
        try {
            for (int r = 0; r < R; ++r) {
                mcrg_chain_desc d{};
                d.L = N;
                d.coupling = J;
                d.seed = mcrg_cpp::next_seed();
                d.n_levels = n_levels;
                d.layout = MCRG_LAYOUT_INT8;
                d.sweeps_per_bin = spb;
                d.max_bins = std::max(1, (n_loc[(size_t)r] + spb - 1) / spb);
                d.sampler = sampler_;
                check(mcrg_chain_create(mcrg_cpp::context(), &d, &chs[(size_t)r]));
                check(mcrg_chain_sweep(chs[(size_t)r], n_samples_eq, 0));
                if (n_loc[(size_t)r] > 0) check(mcrg_chain_sweep(chs[(size_t)r], n_loc[(size_t)r], 1));
            }
            for (int r = 0; r < R; ++r) {
                std::vector<int64_t> rec((size_t)std::max(1, (n_loc[(size_t)r] + spb - 1) / spb) * rs);
                check(mcrg_chain_get_records(chs[(size_t)r], rec.data(), (int64_t)rec.size()));
                for (size_t k = 0; k < rec.size(); ++k) total[k] += rec[k];
            }
        } catch (...) {
            for (mcrg_chain c : chs)
                if (c) mcrg_chain_destroy(c);
            throw;
        }
        for (mcrg_chain c : chs) mcrg_chain_destroy(c);
        return total;
    }
    {
        // replicas advance together in one batch (LDS-resident for 8 <= N <= 512,
        // one HBM-schedule group otherwise)
        std::vector<mcrg_job_desc> jobs((size_t)R);
        for (int r = 0; r < R; ++r) {
            mcrg_job_desc &j = jobs[(size_t)r];
            j.L = N;
            j.coupling = J;
            j.seed = mcrg_cpp::next_seed();
            j.n_therm = n_samples_eq;
            j.n_meas = n_loc[(size_t)r];
            j.n_levels = n_levels;
            j.sweeps_per_bin = spb;
            j.sampler = sampler_;
        }
        mcrg_batch b = nullptr;
        check(mcrg_batch_create(mcrg_cpp::context(), jobs.data(), R, nullptr, &b));
        int32_t left = 1;
        while (left > 0) check(mcrg_batch_run(b, 1 << 30, &left));
        for (int r = 0; r < R; ++r) {
            const int nb = std::max(1, (n_loc[(size_t)r] + spb - 1) / spb);
            std::vector<int64_t> rec((size_t)nb * rs);
            check(mcrg_batch_get_records(b, r, rec.data(), (int64_t)rec.size()));
            for (size_t k = 0; k < rec.size(); ++k) total[k] += rec[k];
        }
        mcrg_batch_destroy(b);
    }
    return total;
}

namespace {
// Weighted straight line D = a + b J through the points with finite weights;
// returns the crossing -a / b (NaN when undefined).
double crossing(const std::vector<double> &J, const std::vector<double> &D, const std::vector<double> &w)
{
    double S = 0, Sx = 0, Sy = 0, Sxx = 0, Sxy = 0;
    for (size_t i = 0; i < J.size(); ++i) {
        S += w[i];
        Sx += w[i] * J[i];
        Sy += w[i] * D[i];
        Sxx += w[i] * J[i] * J[i];
        Sxy += w[i] * J[i] * D[i];
    }
    const double det = S * Sxx - Sx * Sx;
    if (det == 0.0) return std::nan("");
    const double b = (S * Sxy - Sx * Sy) / det, a = (Sy - b * Sx) / S;
    return b != 0.0 ? -a / b : std::nan("");
}

}  // namespace

namespace mcrg_cpp {
namespace {
// Weighted least-squares polynomial of degree deg <= 3 in x = (J - mid) / span
// through the points idx; false when singular.  chi2 = sum w (D - fit)^2.
bool poly_fit(const std::vector<double> &J, const std::vector<double> &D, const std::vector<double> &w,
              const std::vector<size_t> &idx, int deg, double mid, double span, double cf[4], double &chi2)
{
    double M[4][5] = {};
    for (size_t i : idx) {
        const double x = (J[i] - mid) / span;
        const double p[4] = {1.0, x, x * x, x * x * x};
        for (int r = 0; r <= deg; ++r) {
            for (int c = 0; c <= deg; ++c) M[r][c] += w[i] * p[r] * p[c];
            M[r][deg + 1] += w[i] * p[r] * D[i];
        }
    }
    for (int c = 0; c <= deg; ++c) {  // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int r = c + 1; r <= deg; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return false;
        for (int j = 0; j <= deg + 1; ++j) std::swap(M[c][j], M[piv][j]);
        for (int r = 0; r <= deg; ++r) {
            if (r == c) continue;
            const double f = M[r][c] / M[c][c];
            for (int j = 0; j <= deg + 1; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int r = 0; r < 4; ++r) cf[r] = r <= deg ? M[r][deg + 1] / M[r][r] : 0.0;
    chi2 = 0.0;
    for (size_t i : idx) {
        const double x = (J[i] - mid) / span, e = D[i] - (cf[0] + x * (cf[1] + x * (cf[2] + x * cf[3])));
        chi2 += w[i] * e * e;
    }
    return std::isfinite(chi2);
}

// Zero of the polynomial cf in J over [lo, hi] nearest mid (NaN when none).
double poly_zero(const double cf[4], double mid, double span, double lo, double hi)
{
    auto f = [&](double J_) {
        const double x = (J_ - mid) / span;
        return cf[0] + x * (cf[1] + x * (cf[2] + x * cf[3]));
    };
    constexpr int kScan = 512;
    double root = std::nan(""), dist = std::numeric_limits<double>::infinity();
    double pa = lo, fa = f(pa);
    for (int i = 1; i <= kScan; ++i) {
        const double pb = lo + (hi - lo) * i / kScan, fb = f(pb);
        if (fa == 0.0 || fa * fb < 0.0) {
            double a0 = pa, b0 = pb, f0 = fa;
            for (int it = 0; it < 100 && f0 != 0.0 && b0 - a0 > 0.0; ++it) {  // bisection
                const double m = 0.5 * (a0 + b0), fm = f(m);
                if (m <= a0 || m >= b0) break;
                if ((fm < 0.0) == (f0 < 0.0)) {
                    a0 = m;
                    f0 = fm;
                } else {
                    b0 = m;
                }
            }
            const double r = f0 == 0.0 ? a0 : 0.5 * (a0 + b0);
            if (std::fabs(r - mid) < dist) {
                dist = std::fabs(r - mid);
                root = r;
            }
        }
        pa = pb;
        fa = fb;
    }
    if (fa == 0.0 && std::fabs(hi - mid) < dist) root = hi;
    return root;
}
}  // namespace

double zero_crossing(const std::vector<double> &J_in, const std::vector<double> &D_in,
                     const std::vector<double> &w_in)
{
    // points with a positive weight
    std::vector<double> J, D, w;
    for (size_t i = 0; i < std::min(J_in.size(), std::min(D_in.size(), w_in.size())); ++i)
        if (w_in[i] > 0.0 && std::isfinite(w_in[i]) && std::isfinite(D_in[i]) && std::isfinite(J_in[i])) {
            J.push_back(J_in[i]);
            D.push_back(D_in[i]);
            w.push_back(w_in[i]);
        }
    const size_t n = J.size();
    // The weighted straight line through every point whenever it fits (chi2
    // within 3 standard deviations of its mean): in narrow brackets D(J) is
    // straight and noise decides the rest, and the line's zero, unlike a fit
    // confined to the bracket, follows the data out of it.
    if (n >= 3) {
        double S = 0, Sx = 0, Sy = 0, Sxx = 0, Sxy = 0;
        for (size_t i = 0; i < n; ++i) {
            S += w[i];
            Sx += w[i] * J[i];
            Sy += w[i] * D[i];
            Sxx += w[i] * J[i] * J[i];
            Sxy += w[i] * J[i] * D[i];
        }
        const double det = S * Sxx - Sx * Sx;
        if (det != 0.0) {
            const double b = (S * Sxy - Sx * Sy) / det, a = (Sy - b * Sx) / S;
            double chi2 = 0.0;
            for (size_t i = 0; i < n; ++i) chi2 += w[i] * (D[i] - a - b * J[i]) * (D[i] - a - b * J[i]);
            const double dof = (double)n - 2.0;
            if (b != 0.0 && chi2 <= dof + 3.0 * std::sqrt(2.0 * dof)) return -a / b;
        }
    }
    // Otherwise D bends over the bracket: the sign change between neighbours
    // (in J) whose signs are the most significant, |D| sqrt(w) standard errors
    std::vector<size_t> byJ(n);
    for (size_t i = 0; i < n; ++i) byJ[i] = i;
    std::sort(byJ.begin(), byJ.end(), [&](size_t x, size_t y) { return J[x] < J[y]; });
    double best = -1.0, mid = std::nan("");
    for (size_t k = 0; k + 1 < n; ++k) {
        const size_t i = byJ[k], j = byJ[k + 1];
        if (!(D[i] * D[j] <= 0.0) || D[i] == D[j]) continue;
        const double s = std::min(std::fabs(D[i]) * std::sqrt(w[i]), std::fabs(D[j]) * std::sqrt(w[j]));
        if (s > best) {
            best = s;
            mid = J[i] + (J[j] - J[i]) * D[i] / (D[i] - D[j]);
        }
    }
    if (!std::isfinite(mid)) return crossing(J, D, w);
    std::vector<size_t> order(byJ);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return std::fabs(J[x] - mid) < std::fabs(J[y] - mid); });
    // the widest window of nearest points whose weighted quadratic fits
    // (chi2 within 3 standard deviations of its mean) and crosses zero there
    for (size_t k = n; k >= 4; --k) {
        std::vector<size_t> idx(order.begin(), order.begin() + (long)k);
        double lo = J[idx[0]], hi = lo;
        for (size_t i : idx) {
            lo = std::min(lo, J[i]);
            hi = std::max(hi, J[i]);
        }
        double cf[4], chi2;
        if (!(hi > lo) || !poly_fit(J, D, w, idx, 2, mid, hi - lo, cf, chi2)) continue;
        const double dof = (double)k - 3.0;
        if (chi2 > dof + 3.0 * std::sqrt(2.0 * dof)) continue;
        const double r = poly_zero(cf, mid, hi - lo, lo, hi);
        if (std::isfinite(r)) return r;
    }
    // none: the cubic through the four nearest points, else their straight line
    std::vector<size_t> idx(order.begin(), order.begin() + (long)std::min<size_t>(n, 4));
    std::vector<double> Jk, Dk, wk;
    double lo = mid, hi = mid;
    for (size_t i : idx) {
        Jk.push_back(J[i]);
        Dk.push_back(D[i]);
        wk.push_back(w[i]);
        lo = std::min(lo, J[i]);
        hi = std::max(hi, J[i]);
    }
    double cf[4], chi2;
    if (idx.size() == 4 && hi > lo && poly_fit(J, D, w, idx, 3, mid, hi - lo, cf, chi2)) {
        const double r = poly_zero(cf, mid, hi - lo, lo, hi);
        if (std::isfinite(r)) return r;
    }
    const double line = crossing(Jk, Dk, wk);
    return std::isfinite(line) ? line : mid;
}
}  // namespace mcrg_cpp

using mcrg_cpp::zero_crossing;

MonteCarloRenormalizationGroup::Bisection MonteCarloRenormalizationGroup::bisect_critical_point(
    int n_rounds, int n_k, double J_lo, double J_hi, int n_large, int L_large, int n_small, int L_small,
    int n_samples_eq, int n_samples, int sampler_large, int n_samples_eq_small)
{
    if (n_samples_eq_small < 0) n_samples_eq_small = n_samples_eq;
    if (n_rounds < 1 || n_k < 2 || !(J_hi > J_lo) || !(J_lo > 0.0))
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "need n_rounds >= 1, n_k >= 2 and 0 < J_lo < J_hi");
    if (L_small < 8 || L_large <= L_small || (L_large % L_small) != 0 || n_large % n_k || n_small % n_k ||
        n_large < n_k || n_small < n_k || n_samples < 2)
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG,
                              "need L_large = L_small * 2^k, L_small >= 8, and n_large, n_small multiples of n_k");
    int k = 0;
    while ((L_small << k) < L_large) ++k;
    if ((L_small << k) != L_large) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "L_large / L_small must be 2^k");
    // matched levels m = 0 .. n_match-1 (sizes L_small / 2^m, down to 16 sites a side)
    int n_match = 1;
    while (n_match < 4 && (L_small >> n_match) >= 16) ++n_match;
    const int nl_small = n_match - 1, nl_large = k + nl_small, deep = n_match - 1;
    const int rs_l = mcrg_record_size(nl_large), rs_s = mcrg_record_size(nl_small);
    const int per_l = n_large / n_k, per_s = n_small / n_k;
    // jackknife groups of independent chains: replica r of every coupling and
    // size falls in group r % G (chains, unlike bins of one chain, are
    // independent however slowly the sampler decorrelates)
    const int G = std::max(2, std::min(std::min(per_l, per_s), 16));
    if (per_l < 2 || per_s < 2)
        throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "need at least two lattices of each size per coupling");
    Bisection out;
    const auto t0 = std::chrono::steady_clock::now();
    double lo = J_lo, hi = J_hi;
    for (int round = 0; round < n_rounds; ++round) {
        BisectionRound br;
        br.J_lo = lo;
        br.J_hi = hi;
        std::vector<mcrg_job_desc> jobs;
        for (int q = 0; q < n_k; ++q) {
            const double J = lo + (hi - lo) * (q + 0.5) / n_k;
            br.J.push_back(J);
            for (int r = 0; r < per_l; ++r)
                jobs.push_back({L_large, J, mcrg_cpp::next_seed(), n_samples_eq, n_samples, nl_large, n_samples,
                                sampler_large});
            for (int r = 0; r < per_s; ++r)
                jobs.push_back({L_small, J, mcrg_cpp::next_seed(), n_samples_eq_small, n_samples, nl_small, n_samples,
                                MCRG_SAMPLER_METROPOLIS});
        }
        mcrg_batch b = nullptr;
        check(mcrg_batch_create(mcrg_cpp::context(), jobs.data(), (int32_t)jobs.size(), nullptr, &b));
        int32_t left = 1;
        while (left > 0) check(mcrg_batch_run(b, 1 << 30, &left));
        // per coupling and group: NN sums at the matched levels and sample counts
        std::vector<double> SL((size_t)n_k * G * n_match, 0.0), SS((size_t)n_k * G * n_match, 0.0);
        std::vector<double> NL((size_t)n_k * G, 0.0), NS((size_t)n_k * G, 0.0);
        std::vector<int64_t> rec;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const int q = (int)(j / (size_t)(per_l + per_s));
            const int r = (int)(j % (size_t)(per_l + per_s));
            const bool large = r < per_l;
            const int g = (large ? r : r - per_l) % G;
            const int nl = large ? nl_large : nl_small, rs = large ? rs_l : rs_s;
            rec.assign((size_t)rs, 0);
            check(mcrg_batch_get_records(b, (int32_t)j, rec.data(), (int64_t)rec.size()));
            for (int m = 0; m < n_match; ++m) {
                const int lvl = large ? k + m : m;
                (large ? SL : SS)[((size_t)q * G + g) * n_match + m] += (double)rec[(size_t)lvl * 3];
            }
            (large ? NL : NS)[(size_t)q * G + g] += (double)rec[(size_t)(12 * (nl + 1) + 10 * nl + 3)];
            out.site_updates += (double)jobs[j].L * jobs[j].L * (double)(jobs[j].n_therm + jobs[j].n_meas);
        }
        mcrg_batch_destroy(b);
        // D per coupling (NN per site at the matched size), all groups and each group left out
        auto D_of = [&](int q, int m, int skip) {
            double sl = 0, ss = 0, nl_ = 0, ns_ = 0;
            for (int g = 0; g < G; ++g) {
                if (g == skip) continue;
                sl += SL[((size_t)q * G + g) * n_match + m];
                ss += SS[((size_t)q * G + g) * n_match + m];
                nl_ += NL[(size_t)q * G + g];
                ns_ += NS[(size_t)q * G + g];
            }
            const double sites = (double)(L_small >> m) * (double)(L_small >> m);
            return (sl / nl_ - ss / ns_) / sites;
        };
        const double jk = (double)(G - 1) / G;
        std::vector<double> w((size_t)n_k);
        for (int q = 0; q < n_k; ++q) {
            const double d = D_of(q, deep, -1);
            double var = 0.0;
            for (int g = 0; g < G; ++g) var += (D_of(q, deep, g) - d) * (D_of(q, deep, g) - d);
            var *= jk;
            br.D.push_back(d);
            br.D_err.push_back(std::sqrt(var));
            for (int m = 0; m < n_match; ++m) br.D_level.push_back(D_of(q, m, -1));
            w[(size_t)q] = var > 0.0 ? 1.0 / var : 0.0;
        }
        br.Jc = zero_crossing(br.J, br.D, w);
        for (int m = 0; m < n_match; ++m) {
            std::vector<double> Dm((size_t)n_k);
            for (int q = 0; q < n_k; ++q) Dm[(size_t)q] = br.D_level[(size_t)q * n_match + m];
            br.Jc_level.push_back(zero_crossing(br.J, Dm, w));
        }
        double var = 0.0;
        std::vector<double> Dq((size_t)n_k);
        for (int g = 0; g < G; ++g) {
            for (int q = 0; q < n_k; ++q) Dq[(size_t)q] = D_of(q, deep, g);
            const double c = zero_crossing(br.J, Dq, w);
            var += (c - br.Jc) * (c - br.Jc);
        }
        br.Jc_err = std::sqrt(jk * var);
        // progress on stderr, also when quiet
        std::fprintf(stderr, "round %i: J in [%.7f, %.7f] -> Jc = %.7f +- %.7f\n", round + 1, lo, hi, br.Jc,
                     br.Jc_err);
        out.rounds.push_back(br);
        if (!std::isfinite(br.Jc)) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "no crossing: D(J) has no slope");
        // next bracket: a quarter of the width around the crossing, never below
        // three standard errors; a crossing outside the bracket moves it by at
        // most one width without shrinking it
        const double width = hi - lo;
        double centre = br.Jc, half = std::max(3.0 * br.Jc_err, width / 8.0);
        if (br.Jc < lo || br.Jc > hi) {
            centre = std::min(std::max(br.Jc, lo - width / 2.0), hi + width / 2.0);
            half = width / 2.0;
        }
        lo = std::max(1e-6, centre - half);
        hi = centre + half;
    }
    out.Jc = out.rounds.back().Jc;
    out.Jc_err = out.rounds.back().Jc_err;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

void MonteCarloRenormalizationGroup::calc_critical_exponent(int n_samples_eq, int n_samples, int N, double K)
{
    if (!(K <= 0.0)) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "K must be <= 0 (ferromagnetic)");
    const int n_t = levels_ >= 0 ? levels_ : levels_of(N);
    if (n_t < 1) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "need at least one blocking level (N >= 4)");
    if (!quiet_) {
        std::printf("* Calculating critical exponent at K = %lf\n\n", K);
        std::printf("Sampling %i configurations...\n", n_samples);
    }
    const std::string name = dir_ + "/critical_exponent_N_" + std::to_string(N) + "_K_" +
                             mcrg_cpp::rounded_str(K, 7) + ".txt";
    fptr_ = std::fopen(name.c_str(), "w");
    if (!fptr_) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "cannot open " + name);
    std::fprintf(fptr_, "# Number of parallel processes = %i\n", n_processes_);
    std::fprintf(fptr_, "# Number of equilibration samples = %i\n", n_samples_eq);
    std::fprintf(fptr_, "# Number of samples = %i\n", n_samples);
    std::fprintf(fptr_, "# %23s  %25s  %25s\n", "Blocking Level n", "Largest Eigenvalue", "Critical Exponent nu");

    int n_bins = 0;
    const std::vector<int64_t> rec = run_chains(N, K, n_t, n_samples_eq, n_samples, n_bins);
    results_.assign((size_t)n_t, mcrg_rg_result{});
    std::string errs;
    double nu = 0.0;
    for (int n = 0; n < n_t; ++n) {
        mcrg_rg_result r;
        check(mcrg_rg_exponent(rec.data(), n_bins, n_t, n, n_ops_, &r));
        results_[(size_t)n] = r;
        nu = r.nu;
        if (!quiet_) std::printf("n = %i: lambda = %lf, nu = %lf\n", n, r.lambda, r.nu);
        std::fprintf(fptr_, "%25i, %25.10lf, %25.10lf\n", n, r.lambda, r.nu);
    }
    std::fclose(fptr_);
    fptr_ = nullptr;
    // jackknife errors beside the reference-format file, with the integrated
    // autocorrelation time of the NN operator at each level (binning): the
    // jackknife holds while the bins are much longer than it
    FILE *fe = std::fopen((name.substr(0, name.size() - 4) + ".err.txt").c_str(), "w");
    if (fe) {
        std::fprintf(fe, "# delete-one-bin jackknife over %i bins, %i operators\n", n_bins, n_ops_);
        std::fprintf(fe, "# tau_int: integrated autocorrelation time of NN at level n, in sweeps, by binning\n");
        std::fprintf(fe, "# %23s  %25s  %25s  %25s  %25s\n", "Blocking Level n", "nu", "nu error", "y_t error",
                     "tau_int NN");
        for (int n = 0; n < n_t; ++n) {
            mcrg_tau_result tr{};
            const bool have_tau = n_bins >= 2 && mcrg_tau_int(rec.data(), n_bins, n_t, n, 0, &tr) == MCRG_OK;
            std::fprintf(fe, "%25i, %25.10lf, %25.10lf, %25.10lf, %25.4lf\n", n, results_[(size_t)n].nu,
                         results_[(size_t)n].nu_err, results_[(size_t)n].yt_err, have_tau ? tr.tau_int : NAN);
        }
        std::fclose(fe);
    }
    if (!quiet_) std::printf("\n* Critical exponent: nu = %lf\n", nu);
}

double MonteCarloRenormalizationGroup::approx_critical_point(int n_samples_eq, int n_samples, int L, double K)
{
    const int n_t = levels_of(L);
    if (n_t < 1 || L / b_ < 4) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "L must be >= 8 to locate K_c");
    int nb_l = 0, nb_s = 0;
    const int saved_bins = bins_;
    bins_ = 1;
    const std::vector<int64_t> rl = run_chains(L, K, n_t, n_samples_eq, n_samples, nb_l);
    const std::vector<int64_t> rs = run_chains(L / b_, K, n_t - 1, n_samples_eq, n_samples, nb_s);
    bins_ = saved_bins;
    last_kc_.assign((size_t)n_t, 0.0);
    double Kc = K;
    for (int n = 0; n < n_t; ++n) {
        double dK = 0.0;
        check(mcrg_critical_point_step_bins(rl.data(), nb_l, n_t, rs.data(), nb_s, n_t - 1, n, &dK));
        Kc = K + dK;
        last_kc_[(size_t)n] = Kc;
        if (!quiet_) std::printf("n = %i: Kc = %lf\n", n, Kc);
        if (fptr_) {
            std::fprintf(fptr_, "%25i, %25i, %25.10lf, %25.10lf\n", iter_, n, K, Kc);
            std::fflush(fptr_);
        }
    }
    return Kc;
}

double MonteCarloRenormalizationGroup::locate_critical_point(int n_iterations, int n_samples_eq, int n_samples,
                                                              int L, double K0)
{
    if (!(K0 <= 0.0)) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "K0 must be <= 0 (ferromagnetic)");
    if (!quiet_) std::printf("* Locating critical point starting from K0 = %lf\n", K0);
    const std::string name = dir_ + "/critical_point_L_" + std::to_string(L) + "_K_" +
                             mcrg_cpp::rounded_str(K0, 7) + ".txt";
    fptr_ = std::fopen(name.c_str(), "w");
    if (!fptr_) throw mcrg_cpp::Error(MCRG_ERR_INVALID_ARG, "cannot open " + name);
    std::fprintf(fptr_, "# Number of parallel processes = %i\n", n_processes_);
    std::fprintf(fptr_, "# Number of equilibration samples = %i\n", n_samples_eq);
    std::fprintf(fptr_, "# Number of samples = %i\n", n_samples);
    std::fprintf(fptr_, "# %23s  %25s  %25s  %25s\n", "Iteration", "Blocking Level n", "Starting K", "Approximate Kc");
    double K = K0;
    for (iter_ = 1; iter_ <= n_iterations; ++iter_) {
        if (!quiet_) std::printf("\nIteration %i:\n", iter_);
        K = approx_critical_point(n_samples_eq, n_samples, L, K);
    }
    if (!quiet_) {
        std::printf("\n* Critical point: Kc = %lf\n", K);
        std::printf("* Critical temperature: Tc = %lf\n", -1.0 / K);
    }
    std::fclose(fptr_);
    fptr_ = nullptr;
    return K;
}
