// Host engine for the LDS-resident schedule: owns the lattices of many jobs in
// one device arena, groups the active ones into slots and launches chunks.
// Used by the batch and server entry points of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "resident.hpp"

namespace mcrg {

struct JobSpec {
    int L = 0;
    double coupling = 0.0;
    uint64_t seed = 0;
    uint32_t n_therm = 0, n_meas = 0;
    int n_levels = 0;
    int sweeps_per_bin = 1;
    int sampler = 0;  // MCRG_SAMPLER_*
};

// First-fit allocator over one device allocation (offsets in bytes).
class Arena {
  public:
    hipError_t init(size_t bytes);
    void release();
    // Returns false when no free range is large enough.
    bool alloc(size_t bytes, size_t &off);
    void free(size_t off, size_t bytes);
    char *base() const { return base_; }
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }

  private:
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
};

// Measured-sweep time model of one slot on MI355X (from per-L runs of the
// batched workload: about 8 us per sample plus 75 ns per site a lane sweeps).
constexpr uint64_t kSweepFixedNs = 8000;
constexpr uint64_t kSweepPerSiteNs = 75;

// The autotuner keeps the configuration in use unless a candidate's measured
// rate exceeds it by this factor.
constexpr double kTuneMargin = 1.04;

// Longest modelled slot time for which a launch splits slots into parts: a
// tenth of the part wait's 20 s limit.
constexpr uint64_t kPartedSlotCapNs = 2000000000ull;

struct EngineOptions {
    int threads = 256;           // workgroup size of the resident kernel
    int sweeps_per_launch = 250;  // chunk length (the most sweeps any lattice runs per launch)
    int max_group = 256;          // upper bound on lattices per slot
    int min_team = 16;            // lanes per lattice at least (measurement parallelism)
    int max_team = 0;             // lanes per lattice at most (0 = threads): smaller teams
                                  // loop over more words and measure lattices side by side
    int threads_large = 0;        // workgroup size for lattices of L >= large_min_L (0 = threads)
    int large_min_L = 512;
    // LDS a slot of small lattices may take (G halves until it fits);
    // 0 = 32 KiB per 256 threads, which leaves room for five workgroups per CU.
    uint32_t max_slot_lds = 0;
    // Time slicing: when nonzero, a slot runs min(chunk, max(1, quantum_ns /
    // t_sweep)) sweeps per launch, with the measured-sweep time of a slot
    // modelled as t_sweep = kSweepFixedNs + kSweepPerSiteNs * (G L^2 / threads)
    // (a per-sample cost of the measurement and the sites each lane sweeps),
    // so that every workgroup works about quantum_ns; 0 = every slot runs the
    // chunk.
    uint64_t quantum_ns = 0;
    // Parts per slot and launch.  A launch with fewer slots than twice its
    // resident workgroups and more than one per CU gives each workgroup
    // exactly one slot, so the CUs that hold one slot more than the others
    // set the launch time.  Split into sequential parts (queued part-major,
    // each part waiting for the previous one of its slot), the slots' work
    // spreads over the workgroups as they free up.  0 = automatic (up to 4
    // parts of at least 16 sweeps in such launches, else 1).
    int sub_chunks = 0;
};

// Two pinned host buffers used in turn as the source of asynchronous copies:
// a buffer is handed out again only after the event recorded behind its last
// copy has completed, so the host never rewrites memory a copy still reads
// and never waits for the launches queued after that copy.
class PinnedPair {
  public:
    ~PinnedPair();
    // The next buffer, at least `bytes` long (nullptr on error, in `e`).
    void *next(size_t bytes, hipError_t &e);
    // Marks the buffer handed out last as read by copies queued on `s`.
    hipError_t used(hipStream_t s);

  private:
    void *buf_[2] = {nullptr, nullptr};
    size_t cap_[2] = {0, 0};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool pending_[2] = {false, false};
    int cur_ = 1;
};

class ResidentEngine {
  public:
    ~ResidentEngine();
    hipError_t init(int device, size_t arena_bytes, int capacity, const EngineOptions &opt, hipStream_t s);

    // Validates a job for this engine; empty string when it is admissible.
    static std::string check(const JobSpec &j);
    // Device bytes a job needs (spins + records), rounded as the arena rounds.
    static size_t job_bytes(const JobSpec &j);
    // Allocates and initialises the lattice on the stream; returns its id, or
    // -1 when the arena or the descriptor table is full.
    int add(const JobSpec &j);
    void remove(int id);

    // Enqueues one chunk of min(n, sweeps_per_launch) sweeps for every active
    // lattice; `advanced` is the chunk length (0 when nothing is left).
    hipError_t tick(uint32_t n, uint32_t &advanced, struct KernelTimer *tm);
    // Uploads descriptors and draws the start of lattices added since the last call.
    hipError_t flush_pending();

    bool active(int id) const { return id >= 0 && id < (int)jobs_.size() && jobs_[id].live; }
    bool finished(int id) const;
    uint32_t remaining(int id) const;
    uint32_t max_remaining() const;
    int n_live() const;

    hipError_t read_spins(int id, int8_t *host);  // row-major +-1 (synchronous)
    hipError_t read_records(int id, int64_t *host, int64_t n);
    // Records of several lattices through one pinned staging buffer and one wait.
    hipError_t read_records_many(const std::vector<int> &ids, const std::vector<int64_t *> &hosts,
                                 const std::vector<int64_t> &ns);
    // Launches enqueued so far; launch k (0-based) is launches() - 1 right after
    // the tick that enqueued it.
    uint64_t launches() const { return launches_; }
    // Waits for launch k (one of the last two) and what was queued before it.
    hipError_t wait_launch(uint64_t k);
    // read_records_many for lattices that no launch after launch k touches
    // (finished by k or earlier): waits for launch k alone and copies on the
    // engine's readback stream, so later launches keep running meanwhile.
    // Launch k must be one of the last two.
    hipError_t read_records_after(uint64_t k, const std::vector<int> &ids, const std::vector<int64_t *> &hosts,
                                  const std::vector<int64_t> &ns);
    void state(int id, uint32_t &t, uint32_t &n_measured) const;
    const JobSpec &spec(int id) const { return jobs_[id].spec; }
    int n_bins(int id) const { return jobs_[id].n_bins; }
    size_t arena_in_use() const { return arena_.in_use(); }
    size_t arena_capacity() const { return arena_.capacity(); }
    hipStream_t stream() const { return stream_; }
    // Name of the HIP call behind the last error returned by this engine.
    const char *failed_call() const { return failed_; }
    // Autotuning: the next launches each run one candidate configuration
    // (workgroup size, team bounds, slot LDS cap) and are timed with events;
    // then the fastest is kept.  Results do not depend on the configuration,
    // so the jobs advance exactly as they would otherwise.
    void enable_autotune();
    bool tuning() const { return tune_next_ < tune_.size(); }
    const EngineOptions &options() const { return opt_; }
    // Workgroup size of the large-lattice class from the next launch on.
    void set_threads_large(int threads)
    {
        if (threads != opt_.threads_large) {
            opt_.threads_large = threads;
            slots_dirty_ = true;
        }
    }
    // Lattices per slot of side L on NT-thread workgroups.
    int group_size(int L, int nl, int NT) const;
    // Workgroup size of the slots of side L.
    int threads_for(int L) const
    {
        return opt_.threads_large > 0 && L >= opt_.large_min_L ? opt_.threads_large : opt_.threads;
    }
    uint32_t lds_bytes(int L, int G, int nl, int NT) const;
    // LDS bytes one lattice of this shape takes in its slot (slot LDS / G), on
    // NT-thread workgroups or on the ones the engine uses now.
    uint32_t lds_share(int L, int nl, int NT) const;
    uint32_t lds_share(int L, int nl) const { return lds_share(L, nl, threads_for(L)); }

  private:
    struct Job {
        JobSpec spec;
        bool live = false;
        size_t off = 0, bytes = 0;
        int n_bins = 0;
        // host mirror of the device counters (exact: chunks advance deterministically)
        uint32_t t = 0, n_measured = 0, therm_left = 0, meas_left = 0;
    };
    struct TuneCandidate {
        int threads, min_team, threads_large;
        uint32_t max_slot_lds;
        double rate;
    };
    std::vector<TuneCandidate> tune_;
    size_t tune_next_ = 0;
    hipEvent_t tune_a_ = nullptr, tune_b_ = nullptr;
    void rebuild_slots(uint32_t chunk);
    uint32_t slot_sweeps(int L, int G, uint32_t chunk) const;
    // Modelled time of n_sweeps measured sweeps of a slot (the cost model above).
    uint64_t slot_time_ns(int L, int G, uint32_t n_sweeps) const;
    hipError_t launch_classes(KernelTimer *tm);

    // Slots whose LDS allows the same number of workgroups per CU form one
    // launch class; classes launch concurrently on their own streams, each
    // with its own dynamic LDS size, so that a few large lattices do not cut
    // the occupancy of every small one.
    static constexpr int kMaxClasses = 6;
    struct LaunchClass {
        uint32_t first = 0, count = 0, lds = 0;
        int threads = 0, grid = 0;
        uint32_t n_sub = 1;  // parts per slot (sub_chunks)
    };
    std::vector<LaunchClass> classes_;
    hipStream_t aux_[kMaxClasses - 1] = {};
    hipEvent_t ev_fork_ = nullptr;
    hipEvent_t ev_join_[kMaxClasses] = {};
    hipError_t upload_slots();

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    EngineOptions opt_;
    Arena arena_;
    std::vector<Job> jobs_;
    std::vector<int> free_ids_;
    PackedLattice *d_lats_ = nullptr;
    int capacity_ = 0;

    // slot tables: host copies and device mirrors
    std::vector<Slot> slots_;
    std::vector<int32_t> slot_lats_;
    Slot *d_slots_ = nullptr;
    int32_t *d_slot_lats_ = nullptr;
    size_t slot_cap_ = 0, slot_lat_cap_ = 0;
    uint32_t *d_queue_ = nullptr;
    uint32_t *d_done_ = nullptr;  // per-slot progress of parted launches
    size_t done_cap_ = 0;
    uint32_t *d_err_ = nullptr;   // set by a part whose wait timed out
    uint32_t *h_err_ = nullptr;   // its pinned host copy
    int cus_ = 0;
    hipError_t check_parts(hipStream_t s);
    int32_t *d_init_idx_ = nullptr;
    bool slots_dirty_ = true;
    uint32_t slots_chunk_ = 0;
    std::vector<int32_t> pending_init_;
    // admission staging (pinned, in turn) and its device copy; slot-table
    // staging; record readback staging
    PinnedPair stage_;
    PackedLattice *d_stage_ = nullptr;
    PinnedPair slot_stage_;
    int64_t *h_rec_ = nullptr;
    size_t rec_cap_ = 0;
    // one event behind each launch (the last two), the readback stream
    hipEvent_t ev_launch_[2] = {nullptr, nullptr};
    uint64_t launches_ = 0;
    hipStream_t rb_ = nullptr;
    hipError_t copy_records(hipStream_t s, const std::vector<int> &ids, const std::vector<int64_t *> &hosts,
                            const std::vector<int64_t> &ns);
    const char *failed_ = "";
};

}  // namespace mcrg
