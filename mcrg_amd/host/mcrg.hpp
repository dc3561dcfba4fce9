// C++ interface of the MCRG hot path with the reference's class and method
// names (kim-jane/MCRG src/lattice.hpp, src/ising.hpp, src/mcrg.hpp), running on
// the MI355X through the C ABI of libmcrg_amd.so (include/mcrg_abi.h).
//
// Differences from the reference, by design:
//  * One sample (sample_new_configuration) is one red/black checkerboard
//    Metropolis sweep (the reference flips one Wolff cluster).
//  * Seeds are reproducible: every new chain takes the next value of a
//    process-wide Philox seed sequence (mcrg_cpp::set_seed) instead of
//    std::random_device.
//  * No MPI.  The reference's ranks become replicas on the one device
//    (MonteCarloRenormalizationGroup::set_replicas); samples are split over them
//    with split_samples' rule and every replica equilibrates on its own.
//  * imat / vec2D are small value types of this file instead of Eigen types.
//  * calc_magnetization divides in floating point (the reference's integer
//    division returns 0 or +-1, src/ising.cpp:178).
//  * block_spin_transformation is fused into the device measurement; no blocked
//    Lattice objects are built per sample.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcrg_abi.h"

// Column-major int matrix (the role of Eigen's imat, src/definitions.hpp:16).
class imat {
  public:
    imat() = default;
    imat(int rows, int cols) { resize(rows, cols); }
    void resize(int rows, int cols)
    {
        r_ = rows;
        c_ = cols;
        v_.assign((size_t)rows * (size_t)cols, 0);
    }
    int rows() const { return r_; }
    int cols() const { return c_; }
    int &operator()(int i, int j) { return v_[(size_t)j * (size_t)r_ + (size_t)i]; }
    int operator()(int i, int j) const { return v_[(size_t)j * (size_t)r_ + (size_t)i]; }
    long long sum() const
    {
        long long s = 0;
        for (int x : v_) s += x;
        return s;
    }
    int *data() { return v_.data(); }
    const int *data() const { return v_.data(); }

  private:
    int r_ = 0, c_ = 0;
    std::vector<int> v_;
};

// Two-component vector (the role of Eigen's vec2D).
struct vec2D {
    double v[2] = {0.0, 0.0};
    double &operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
};

namespace mcrg_cpp {
// Thrown for any non-OK status of the C ABI (message from mcrg_last_error()).
class Error : public std::runtime_error {
  public:
    Error(int status, const std::string &what) : std::runtime_error(what), status(status) {}
    int status;
};
void check(mcrg_status st);
// Device used by every object of this process (default 0), set before the first object.
void set_device(int device);
mcrg_ctx context();
// Start of the seed sequence (default 1); chain k of the process gets seed
// splitmix64(base + k).
void set_seed(uint64_t base);
uint64_t next_seed();
std::string rounded_str(double num, int precision);  // get_rounded_str, src/definitions.cpp:70-77
bool write_iter(int i);                                // src/definitions.cpp:46-68
int split_samples(int rank, int n_processes, int n_samples);  // src/definitions.cpp:79-87
// Zero of D(J) sampled at couplings J with weights w (1 / variance), as the
// K-bisection locates it.  The weighted straight line through every point when
// it fits (chi^2 within three standard deviations of its mean): its zero may
// lie outside the bracket, so a noisy narrow bracket does not pin the answer
// to its own centre.  Over a wide bracket D(J) bends (it saturates on either
// side of the critical coupling) and the line is rejected; then, among the
// points with positive weight: the sign change between neighbours whose signs
// are the most significant (|D| sqrt(w)), the widest window of points nearest
// it whose weighted quadratic fits and crosses zero inside the window, that
// zero nearest the sign change; failing that, the cubic through the four
// nearest points.  Without a sign change: the straight line, an
// extrapolation.  NaN when undefined.
double zero_crossing(const std::vector<double> &J, const std::vector<double> &D, const std::vector<double> &w);
}  // namespace mcrg_cpp

class Lattice {
  public:
    explicit Lattice(int N);     // random +-1 start (src/lattice.cpp:3-15)
    Lattice(int a, imat spins);  // given spins, lattice spacing a (src/lattice.cpp:18-30)
    ~Lattice();
    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;

    int N_;
    int n_spins_;
    int a_;
    imat spins_;  // host view of the device spins, refreshed by sync_spins()

    void write_spins(FILE *fptr);
    void display_spins();
    int choose_random_spin();
    double calc_nearest_neighbor_interaction();  // double-counted NN, as the reference
    vec2D calc_interactions();                   // double-counted NN, NNN
    imat nearest_neighbors(int i, int j);
    imat next_nearest_neighbors(int i, int j);

    // additions
    void sync_spins();  // device -> spins_
    void push_spins();  // spins_ -> device
    mcrg_chain chain() const { return chain_; }
    uint64_t seed() const { return seed_; }
    void set_coupling(double J);  // J = -K; no-op when unchanged

  private:
    mcrg_chain chain_ = nullptr;
    double J_ = 0.0;
    uint64_t seed_ = 0;
    uint64_t pick_ = 0;
};

class IsingModel {
  public:
    explicit IsingModel(double K);  // K < 0 ferromagnetic, as the reference
    ~IsingModel() = default;

    void equilibrate(std::shared_ptr<Lattice> pLattice, int n_samples_eq, bool write);
    void sample_new_configuration(std::shared_ptr<Lattice> pLattice);
    double calc_magnetization(std::shared_ptr<Lattice> pLattice);
    double calc_energy(std::shared_ptr<Lattice> pLattice);

  private:
    double K_;
    void bind(Lattice &lat);
};

class MonteCarloRenormalizationGroup {
  public:
    explicit MonteCarloRenormalizationGroup(int b);
    ~MonteCarloRenormalizationGroup() = default;

    int b_;

    void calc_critical_exponent(int n_samples_eq, int n_samples, int N, double K);
    double locate_critical_point(int n_iterations, int n_samples_eq, int n_samples, int L, double K0);

    // additions
    void set_operators(int n_ops);    // 2 (NN, NNN: the reference) or 3 (+ plaquette)
    void set_replicas(int replicas);  // the reference's MPI ranks (default 1)
    void set_bins(int bins);          // jackknife bins (default 50)
    void set_levels(int levels);      // blockings; default floor(log2 N) - 1
    void set_output_dir(const std::string &dir);
    void set_quiet(bool quiet);
    void set_sampler(int sampler);    // MCRG_SAMPLER_* for calc_critical_exponent and the locator
    // Storage of the replicas' lattices: -1 = automatic (bit-packed in LDS for
    // Metropolis at 8 <= N <= 512, int8 in device memory otherwise);
    // MCRG_LAYOUT_INT8 = one int8 chain per replica (mcrg_chain_*);
    // MCRG_LAYOUT_PACKED = the bit-packed schedule, an error where it cannot
    // run.  Results are bit-identical under every layout.
    void set_layout(int layout);
    // exponents of the last calc_critical_exponent, one per level
    const std::vector<mcrg_rg_result> &results() const { return results_; }
    // per-level estimates of the last approx_critical_point
    const std::vector<double> &last_kc() const { return last_kc_; }

    // K-bisection for the critical coupling with lattices of two sizes run
    // concurrently (no reference counterpart; it batches the matching idea of
    // approx_critical_point, src/mcrg.cpp:191-310).  Each round runs, for n_k
    // couplings J spanning [J_lo, J_hi], n_large / n_k lattices of L_large and
    // n_small / n_k of L_small in one batch, and compares the nearest-neighbour
    // correlation per site of L_large blocked k + m times with L_small blocked
    // m times (k = log2(L_large / L_small); matched levels m = 0..3 while
    // L_small / 2^m >= 16: same final size).  D(J) = large - small at the
    // deepest matched level changes sign at the critical coupling; a local fit
    // around the sign change (mcrg_cpp::zero_crossing) gives the crossing J0
    // and its jackknife error over groups of independent chains, and the next
    // bracket is
    // J0 +- max(3 err, width / 8) (a crossing outside the bracket moves it by at
    // most one width).  J = |K|.
    struct BisectionRound {
        double J_lo = 0, J_hi = 0, Jc = 0, Jc_err = 0;
        std::vector<double> J, D, D_err;
        std::vector<double> D_level;   // [coupling][matched level m]
        std::vector<double> Jc_level;  // crossing at each matched level
    };
    struct Bisection {
        std::vector<BisectionRound> rounds;
        double Jc = 0, Jc_err = 0, seconds = 0, site_updates = 0;
    };
    // The large lattices use `sampler_large` (Swendsen-Wang by default: they
    // equilibrate in tens of sweeps where Metropolis needs ~L^2.2); the small
    // ones run Metropolis on the LDS-resident schedule with n_samples_eq_small
    // equilibration sweeps (-1: n_samples_eq).
    Bisection bisect_critical_point(int n_rounds, int n_k, double J_lo, double J_hi, int n_large, int L_large,
                                    int n_small, int L_small, int n_samples_eq, int n_samples,
                                    int sampler_large = MCRG_SAMPLER_SWENDSEN_WANG, int n_samples_eq_small = -1);

  private:
    double approx_critical_point(int n_samples_eq, int n_samples, int L, double K);
    // records of `replicas` chains of side N at coupling K (summed per bin)
    std::vector<int64_t> run_chains(int N, double K, int n_levels, int n_samples_eq, int n_samples, int &n_bins);

    int n_processes_ = 1;
    int n_ops_ = 2;
    int sampler_ = MCRG_SAMPLER_METROPOLIS;
    int layout_ = -1;
    int bins_ = 50;
    int levels_ = -1;
    int iter_ = 0;
    bool quiet_ = false;
    std::string dir_ = ".";
    FILE *fptr_ = nullptr;
    std::vector<mcrg_rg_result> results_;
    std::vector<double> last_kc_;
};
