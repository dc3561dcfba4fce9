// HBM-schedule host runtime: device blocks, sweep runs with replayed graphs,
// lockstep lattice groups.
#include "hbm_group.hpp"

#include <cstring>

namespace mcrg {

static size_t align256(size_t v) { return (v + 255u) & ~size_t(255u); }

HbmBlock hbm_block(int L, int n_levels, int max_bins, int sampler)
{
    HbmBlock b;
    const size_t n_spins = (size_t)L * (size_t)L;
    // levels 1 .. max(n_levels, 1): level 1 is always written by the level-0 measurement
    const uint32_t top = n_levels > 1 ? (uint32_t)n_levels : 1u;
    const size_t n_lvl = L >= 2 ? level_offset((uint32_t)L, top + 1u) : 0u;
    b.spins = 0;
    b.lvl = align256(b.spins + n_spins);
    b.sums = align256(b.lvl + n_lvl * 4u);
    b.state = align256(b.sums + kSumWords * 8u);
    b.rec = align256(b.state + 2u * 4u);
    b.labels = align256(b.rec + (size_t)max_bins * record_size((uint32_t)n_levels) * 8u);
    b.bytes = align256(b.labels + (sampler == MCRG_SAMPLER_SWENDSEN_WANG ? n_spins * 4u : 0u));
    return b;
}

HbmLattice hbm_descriptor(char *base, const HbmBlock &b, int L, int n_levels, int sweeps_per_bin, uint64_t seed,
                          double coupling, int sampler)
{
    HbmLattice h;
    std::memset(&h, 0, sizeof(h));
    h.spins = reinterpret_cast<int8_t *>(base + b.spins);
    h.lvl = reinterpret_cast<uint32_t *>(base + b.lvl);
    h.sums = reinterpret_cast<unsigned long long *>(base + b.sums);
    h.state = reinterpret_cast<uint32_t *>(base + b.state);
    h.records = reinterpret_cast<int64_t *>(base + b.rec);
    h.k0 = (uint32_t)seed;
    h.k1 = (uint32_t)(seed >> 32);
    h.acc = make_accept(coupling);
    h.L = L;
    h.n_levels = n_levels;
    h.sweeps_per_bin = sweeps_per_bin;
    h.sampler = sampler;
    h.labels = sampler == MCRG_SAMPLER_SWENDSEN_WANG ? reinterpret_cast<int32_t *>(base + b.labels) : nullptr;
    h.bond_thr = bond_threshold(coupling);
    return h;
}

hipError_t enqueue_sweeps(const HbmLattice *d, int n_lats, int L, int n_levels, int sampler, uint32_t n, bool measure,
                          hipStream_t s, KernelTimer *tm)
{
    hipError_t e;
    for (uint32_t i = 0; i < n; ++i) {
        if (sampler == MCRG_SAMPLER_SWENDSEN_WANG) {
            const int ev = tm ? tm->begin(kKindSweep, s) : -1;
            if ((e = launch_sw_sweep(d, n_lats, L, i, s)) != hipSuccess) return e;
            if (tm) tm->end(ev, s);
        } else {
            for (uint32_t c = 0; c < 2u; ++c) {
                const int ev = tm ? tm->begin(kKindSweep, s) : -1;
                if ((e = launch_int8_half_sweep(d, n_lats, L, i, c, s)) != hipSuccess) return e;
                if (tm) tm->end(ev, s);
            }
        }
        if (measure) {
            int ev = tm ? tm->begin(kKindMeasure0, s) : -1;
            if ((e = launch_int8_measure0(d, n_lats, L, i, s)) != hipSuccess) return e;
            if (tm) tm->end(ev, s);
            ev = tm ? tm->begin(kKindLevels, s) : -1;
            if ((e = launch_levels_accumulate(d, n_lats, L, n_levels, i, i, s)) != hipSuccess) return e;
            if (tm) tm->end(ev, s);
        }
    }
    return hipSuccess;
}

SweepRunner::~SweepRunner()
{
    for (hipGraphExec_t &g : exec_)
        if (g) (void)hipGraphExecDestroy(g);
}

hipError_t SweepRunner::run(const HbmLattice *d, int n_lats, int L, int n_levels, int sampler, uint32_t n, bool measure,
                            hipStream_t s, KernelTimer *tm)
{
    hipError_t e;
    uint32_t done = 0;
    if (!tm && n >= kGraphSweeps) {
        hipGraphExec_t &exec = exec_[measure ? 1 : 0];
        if (!exec) {
            // kGraphSweeps sweeps and the counter advance, captured once: the
            // kernels read the sweep and sample counters from device memory,
            // so every replay continues where the previous one stopped
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
            hipError_t ce = enqueue_sweeps(d, n_lats, L, n_levels, sampler, kGraphSweeps, measure, s, nullptr);
            if (ce == hipSuccess) ce = launch_advance(d, n_lats, kGraphSweeps, measure ? kGraphSweeps : 0u, s);
            e = hipStreamEndCapture(s, &g);
            if (ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return ce;
            }
            if (e != hipSuccess) return e;
            e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                exec = nullptr;
                return e;
            }
        }
        for (; done + kGraphSweeps <= n; done += kGraphSweeps)
            if ((e = hipGraphLaunch(exec, s)) != hipSuccess) return e;
    }
    const uint32_t rest = n - done;
    if (rest == 0) return hipSuccess;
    if ((e = enqueue_sweeps(d, n_lats, L, n_levels, sampler, rest, measure, s, tm)) != hipSuccess) return e;
    return launch_advance(d, n_lats, rest, measure ? rest : 0u, s);
}

// ------------------------------------------------------------------ groups
HbmGroup::~HbmGroup()
{
    if (d_block_) (void)hipFree(d_block_);
}

bool HbmGroup::same_shape(const JobSpec &a, const JobSpec &b)
{
    return a.L == b.L && a.n_levels == b.n_levels && a.n_therm == b.n_therm && a.n_meas == b.n_meas &&
           a.sweeps_per_bin == b.sweeps_per_bin && a.sampler == b.sampler;
}

std::string HbmGroup::check(const JobSpec &j)
{
    if (!is_pow2(j.L) || j.L < 2 || j.L > 16384) return "L must be a power of two in [2, 16384]";
    if (!(j.coupling >= 0.0) || !std::isfinite(j.coupling)) return "coupling J = |K| must be >= 0";
    if (j.n_levels < 0 || j.n_levels > ilog2(j.L) - 1 || j.n_levels > (int)kMaxLevels)
        return "n_levels must be in [0, log2(L) - 1]";
    if (j.sweeps_per_bin < 1) return "sweeps_per_bin must be >= 1";
    if (!bin_fits(j.L, j.sweeps_per_bin)) return MCRG_BIN_LIMIT_MSG;
    if ((uint64_t)j.n_therm + (uint64_t)j.n_meas > 0xFFFFFFFFull) return "too many sweeps";
    if (j.sampler != MCRG_SAMPLER_METROPOLIS && j.sampler != MCRG_SAMPLER_SWENDSEN_WANG)
        return "sampler must be MCRG_SAMPLER_METROPOLIS or MCRG_SAMPLER_SWENDSEN_WANG";
    return std::string();
}

hipError_t HbmGroup::init(const std::vector<JobSpec> &members, hipStream_t s)
{
    members_ = members;
    const JobSpec &j = members_[0];
    n_bins_ = (int)((j.n_meas + (uint32_t)j.sweeps_per_bin - 1u) / (uint32_t)j.sweeps_per_bin);
    if (n_bins_ < 1) n_bins_ = 1;
    therm_left_ = j.n_therm;
    meas_left_ = j.n_meas;
    blk_ = hbm_block(j.L, j.n_levels, n_bins_, j.sampler);
    const size_t n = members_.size();
    const size_t desc_off = blk_.bytes * n;
    hipError_t e;
    void *p = nullptr;
    if ((e = hipMalloc(&p, desc_off + sizeof(HbmLattice) * n)) != hipSuccess) return e;
    d_block_ = static_cast<char *>(p);
    d_desc_ = reinterpret_cast<HbmLattice *>(d_block_ + desc_off);
    host_.resize(n);
    for (size_t k = 0; k < n; ++k)
        host_[k] = hbm_descriptor(d_block_ + k * blk_.bytes, blk_, j.L, j.n_levels, j.sweeps_per_bin,
                                  members_[k].seed, members_[k].coupling, j.sampler);
    if ((e = hipMemsetAsync(d_block_, 0, desc_off, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(d_desc_, host_.data(), sizeof(HbmLattice) * n, hipMemcpyHostToDevice, s)) !=
        hipSuccess)
        return e;
    if ((e = launch_int8_init(d_desc_, (int)n, j.L, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);  // host_ is read by the copy above
}

hipError_t HbmGroup::advance(uint32_t n, hipStream_t s, KernelTimer *tm)
{
    hipError_t e;
    const uint32_t therm = std::min(n, therm_left_);
    if (therm > 0) {
        if ((e = runner_.run(d_desc_, size(), spec().L, spec().n_levels, spec().sampler, therm, false, s, tm)) != hipSuccess) return e;
        therm_left_ -= therm;
        t_ += therm;
    }
    const uint32_t meas = std::min(n - therm, meas_left_);
    if (meas > 0) {
        if ((e = runner_.run(d_desc_, size(), spec().L, spec().n_levels, spec().sampler, meas, true, s, tm)) != hipSuccess) return e;
        meas_left_ -= meas;
        t_ += meas;
        q_ += meas;
    }
    return hipSuccess;
}

hipError_t HbmGroup::read_records(int member, int64_t *host, int64_t n, hipStream_t s)
{
    hipError_t e;
    if ((e = hipMemcpyAsync(host, host_[(size_t)member].records, (size_t)n * 8u, hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
        return e;
    return hipStreamSynchronize(s);
}

hipError_t HbmGroup::read_spins(int member, int8_t *host, hipStream_t s)
{
    hipError_t e;
    const size_t n = (size_t)spec().L * (size_t)spec().L;
    if ((e = hipMemcpyAsync(host, host_[(size_t)member].spins, n, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

}  // namespace mcrg
