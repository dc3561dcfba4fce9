// Host side of the HBM (int8) schedule: per-lattice device blocks, sweep
// enqueueing with replayed HIP graphs, and groups of lattices that advance in
// lockstep through shared launches.  Used by the chain and batch entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"
#include "hbm_kernels.hpp"
#include "runtime.hpp"

namespace mcrg {

// Byte offsets (256-aligned) of one lattice's buffers inside its device block.
struct HbmBlock {
    size_t spins = 0, lvl = 0, sums = 0, state = 0, rec = 0, labels = 0, bytes = 0;
};
HbmBlock hbm_block(int L, int n_levels, int max_bins, int sampler);
// Descriptor of the lattice whose block starts at `base`.
HbmLattice hbm_descriptor(char *base, const HbmBlock &b, int L, int n_levels, int sweeps_per_bin, uint64_t seed,
                          double coupling, int sampler);

// Launches n sweeps (t_off 0..n-1) of n_lats lattices of side L without
// advancing their counters.
hipError_t enqueue_sweeps(const HbmLattice *d, int n_lats, int L, int n_levels, int sampler, uint32_t n, bool measure,
                          hipStream_t s, KernelTimer *tm);

// Runs of sweeps followed by the counter advance.  Whole blocks of
// kGraphSweeps sweeps replay one captured graph per (descriptors, measure);
// the rest, and every sweep while kernel timing is on, launch eagerly.
class SweepRunner {
  public:
    static constexpr uint32_t kGraphSweeps = 32;
    SweepRunner() = default;
    SweepRunner(const SweepRunner &) = delete;
    SweepRunner &operator=(const SweepRunner &) = delete;
    ~SweepRunner();
    hipError_t run(const HbmLattice *d, int n_lats, int L, int n_levels, int sampler, uint32_t n, bool measure, hipStream_t s,
                   KernelTimer *tm);

  private:
    hipGraphExec_t exec_[2] = {nullptr, nullptr};
};

// Lattices that share L, level count, sweep counts and bin size, advanced
// together (blockIdx.y = member) on the HBM schedule.
class HbmGroup {
  public:
    HbmGroup() = default;
    HbmGroup(const HbmGroup &) = delete;
    HbmGroup &operator=(const HbmGroup &) = delete;
    ~HbmGroup();
    // Same-shaped jobs (L, n_levels, n_therm, n_meas, sweeps_per_bin, sampler equal).
    static bool same_shape(const JobSpec &a, const JobSpec &b);
    static std::string check(const JobSpec &j);
    hipError_t init(const std::vector<JobSpec> &members, hipStream_t s);
    // Advances every member by min(n, remaining) sweeps (asynchronous).
    hipError_t advance(uint32_t n, hipStream_t s, KernelTimer *tm);
    uint32_t remaining() const { return therm_left_ + meas_left_; }
    int size() const { return (int)members_.size(); }
    int n_bins() const { return n_bins_; }
    const JobSpec &spec() const { return members_[0]; }
    void state(uint32_t &t, uint32_t &n_measured) const
    {
        t = t_;
        n_measured = q_;
    }
    hipError_t read_records(int member, int64_t *host, int64_t n, hipStream_t s);
    hipError_t read_spins(int member, int8_t *host, hipStream_t s);

  private:
    std::vector<JobSpec> members_;
    HbmBlock blk_;
    char *d_block_ = nullptr;
    HbmLattice *d_desc_ = nullptr;
    std::vector<HbmLattice> host_;
    int n_bins_ = 1;
    uint32_t t_ = 0, q_ = 0, therm_left_ = 0, meas_left_ = 0;
    SweepRunner runner_;
};

}  // namespace mcrg
