// Host engine of the LDS-resident schedule.
#include "engine.hpp"

#include <algorithm>
#include <cstring>

#include "runtime.hpp"

namespace mcrg {

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ------------------------------------------------------------------ Arena
hipError_t Arena::init(size_t bytes)
{
    cap_ = round_up(bytes, 256);
    used_ = 0;
    free_.clear();
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, cap_);
    if (e != hipSuccess) {
        cap_ = 0;
        return e;
    }
    base_ = static_cast<char *>(p);
    free_[0] = cap_;
    return hipSuccess;
}

void Arena::release()
{
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = used_ = 0;
    free_.clear();
}

bool Arena::alloc(size_t bytes, size_t &off)
{
    bytes = round_up(bytes, 256);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= bytes) {
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            return true;
        }
    }
    return false;
}

void Arena::free(size_t off, size_t bytes)
{
    bytes = round_up(bytes, 256);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// ------------------------------------------------------------------ PinnedPair
PinnedPair::~PinnedPair()
{
    for (int i = 0; i < 2; ++i) {
        if (ev_[i]) {
            (void)hipEventSynchronize(ev_[i]);
            (void)hipEventDestroy(ev_[i]);
        }
        if (buf_[i]) (void)hipHostFree(buf_[i]);
    }
}

void *PinnedPair::next(size_t bytes, hipError_t &e)
{
    e = hipSuccess;
    const int i = cur_ ^ 1;
    if (!ev_[i] && (e = hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming)) != hipSuccess) return nullptr;
    // the copies that read this buffer last have to be done before it is rewritten
    if (pending_[i] && (e = hipEventSynchronize(ev_[i])) != hipSuccess) return nullptr;
    pending_[i] = false;
    if (bytes > cap_[i]) {
        if (buf_[i]) (void)hipHostFree(buf_[i]);
        buf_[i] = nullptr;
        cap_[i] = 0;
        const size_t cap = std::max<size_t>(bytes, 1u << 16);
        if ((e = hipHostMalloc(&buf_[i], cap, hipHostMallocDefault)) != hipSuccess) return nullptr;
        cap_[i] = cap;
    }
    cur_ = i;
    return buf_[i];
}

hipError_t PinnedPair::used(hipStream_t s)
{
    pending_[cur_] = true;
    return hipEventRecord(ev_[cur_], s);
}

// ------------------------------------------------------------------ engine
ResidentEngine::~ResidentEngine()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (hipStream_t a : aux_)
        if (a) {
            (void)hipStreamSynchronize(a);
            (void)hipStreamDestroy(a);
        }
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (tune_a_) (void)hipEventDestroy(tune_a_);
    if (tune_b_) (void)hipEventDestroy(tune_b_);
    for (hipEvent_t e : ev_join_)
        if (e) (void)hipEventDestroy(e);
    arena_.release();
    if (d_lats_) (void)hipFree(d_lats_);
    if (d_slots_) (void)hipFree(d_slots_);
    if (d_slot_lats_) (void)hipFree(d_slot_lats_);
    if (d_queue_) (void)hipFree(d_queue_);
    if (d_done_) (void)hipFree(d_done_);
    if (d_err_) (void)hipFree(d_err_);
    if (d_init_idx_) (void)hipFree(d_init_idx_);
    if (d_stage_) (void)hipFree(d_stage_);
    if (h_rec_) (void)hipHostFree(h_rec_);
    if (h_err_) (void)hipHostFree(h_err_);
    for (hipEvent_t e : ev_launch_)
        if (e) (void)hipEventDestroy(e);
    if (rb_) {
        (void)hipStreamSynchronize(rb_);
        (void)hipStreamDestroy(rb_);
    }
}

hipError_t ResidentEngine::init(int device, size_t arena_bytes, int capacity, const EngineOptions &opt,
                                hipStream_t s)
{
    device_ = device;
    stream_ = s;
    opt_ = opt;
    capacity_ = capacity;
    hipError_t e;
    if ((e = arena_.init(arena_bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_lats_, sizeof(PackedLattice) * (size_t)capacity)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_queue_, 64u * kMaxClasses)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_err_, sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = hipMemset(d_err_, 0, sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_init_idx_, sizeof(int32_t) * (size_t)capacity)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_stage_, sizeof(PackedLattice) * (size_t)capacity)) != hipSuccess) return e;
    for (hipStream_t &a : aux_)
        if ((e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming)) != hipSuccess) return e;
    for (hipEvent_t &ev : ev_join_)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
    for (hipEvent_t &ev : ev_launch_)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipStreamCreateWithFlags(&rb_, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&h_err_), sizeof(uint32_t), hipHostMallocDefault)) !=
        hipSuccess)
        return e;
    return hipSuccess;
}

std::string ResidentEngine::check(const JobSpec &j)
{
    if (!is_pow2(j.L) || j.L < kResidentMinL || j.L > kResidentMaxL)
        return "L must be a power of two in [8, 512] for the LDS-resident schedule";
    if (!(j.coupling >= 0.0) || !std::isfinite(j.coupling)) return "coupling J = |K| must be >= 0";
    if (j.n_levels < 0 || j.n_levels > ilog2(j.L) - 1) return "n_levels must be in [0, log2(L) - 1]";
    if (j.sweeps_per_bin < 1) return "sweeps_per_bin must be >= 1";
    if (!bin_fits(j.L, j.sweeps_per_bin)) return MCRG_BIN_LIMIT_MSG;
    if ((uint64_t)j.n_therm + (uint64_t)j.n_meas > 0xFFFFFFFFull) return "too many sweeps";
    if (j.sampler != MCRG_SAMPLER_METROPOLIS) return "the LDS-resident schedule runs the Metropolis sampler only";
    return std::string();
}

static int bins_of(const JobSpec &j)
{
    const int b = (int)((j.n_meas + (uint32_t)j.sweeps_per_bin - 1) / (uint32_t)j.sweeps_per_bin);
    return b > 0 ? b : 1;
}

size_t ResidentEngine::job_bytes(const JobSpec &j)
{
    const size_t spins = round_up(2u * cb_words((uint32_t)j.L) * 4u, 256);
    const size_t rec = round_up((size_t)bins_of(j) * record_size((uint32_t)j.n_levels) * 8u, 256);
    return spins + rec;
}

uint32_t ResidentEngine::lds_bytes(int L, int G, int nl, int NT) const
{
    const uint32_t T = (uint32_t)NT / (uint32_t)G;
    return slot_layout((uint32_t)L, (uint32_t)G, (uint32_t)nl, T).total * 4u;
}

void ResidentEngine::enable_autotune()
{
    // the first entry repeats the configuration in use to absorb first-launch costs
    const EngineOptions o = opt_;
    tune_ = {{o.threads, o.min_team, o.threads_large, o.max_slot_lds, 0.0},
             {o.threads, o.min_team, o.threads_large, o.max_slot_lds, 0.0},
             {128, 16, 0, 0, 0.0},
             {256, 16, 0, 0, 0.0},
             {512, 16, 0, 0, 0.0},
             {256, 32, 0, 0, 0.0},
             {256, 8, 0, 0, 0.0},
             {256, 16, 512, 0, 0.0},
             {256, 16, 0, 65536, 0.0}};
    tune_next_ = 0;
}

int ResidentEngine::group_size(int L, int nl, int NT) const
{
    const int W = (int)cb_words((uint32_t)L);
    int G = W >= NT ? 1 : NT / W;
    if (opt_.max_team > 0) G = std::max(G, NT / opt_.max_team);
    int cap = 1;
    while (cap * 2 <= opt_.max_group) cap *= 2;  // positions are a power of two
    G = std::min(G, cap);
    // a team of at least min_team lanes shares each lattice's measurement, but
    // no more than max(4, W) lanes: a lattice of W words per colour keeps the
    // rest of a larger team idle in its sweeps (L = 8: 121 -> 300 updates/ns
    // with 4 lanes instead of 16; L = 16: 372 -> 574)
    const int team_min = std::min(std::max(1, opt_.min_team), std::max(4, W));
    G = std::min(G, std::max(1, NT / team_min));
    // keep a slot's LDS modest so that several workgroups share a CU
    const uint32_t soft = opt_.max_slot_lds ? opt_.max_slot_lds : 32768u * (uint32_t)NT / 256u;
    while (G > 1 && lds_bytes(L, G, nl, NT) > soft) G /= 2;
    return std::max(G, 1);
}

uint32_t ResidentEngine::lds_share(int L, int nl, int NT) const
{
    const int G = group_size(L, nl, NT);
    return (lds_bytes(L, G, nl, NT) + (uint32_t)G - 1u) / (uint32_t)G;
}

uint64_t ResidentEngine::slot_time_ns(int L, int G, uint32_t n_sweeps) const
{
    const uint64_t sites_per_lane = (uint64_t)G * (uint64_t)L * (uint64_t)L / (uint64_t)threads_for(L);
    return (uint64_t)n_sweeps * (kSweepFixedNs + kSweepPerSiteNs * sites_per_lane);
}

uint32_t ResidentEngine::slot_sweeps(int L, int G, uint32_t chunk) const
{
    if (opt_.quantum_ns == 0) return chunk;
    const uint64_t sites_per_lane = (uint64_t)G * (uint64_t)L * (uint64_t)L / (uint64_t)threads_for(L);
    const uint64_t n = opt_.quantum_ns / (kSweepFixedNs + kSweepPerSiteNs * sites_per_lane);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n, chunk));
}

int ResidentEngine::add(const JobSpec &j)
{
    int id;
    if (!free_ids_.empty()) {
        id = free_ids_.back();
    } else {
        if ((int)jobs_.size() >= capacity_) return -1;
        id = (int)jobs_.size();
    }
    const size_t bytes = job_bytes(j);
    size_t off = 0;
    if (!arena_.alloc(bytes, off)) return -1;
    if (!free_ids_.empty()) free_ids_.pop_back();
    else jobs_.emplace_back();
    Job &job = jobs_[(size_t)id];
    job = Job();
    job.spec = j;
    job.live = true;
    job.off = off;
    job.bytes = bytes;
    job.n_bins = bins_of(j);
    job.therm_left = j.n_therm;
    job.meas_left = j.n_meas;
    pending_init_.push_back(id);
    slots_dirty_ = true;
    return id;
}

void ResidentEngine::remove(int id)
{
    if (!active(id)) return;
    Job &job = jobs_[(size_t)id];
    arena_.free(job.off, job.bytes);
    job.live = false;
    free_ids_.push_back(id);
    pending_init_.erase(std::remove(pending_init_.begin(), pending_init_.end(), id), pending_init_.end());
    slots_dirty_ = true;
}

bool ResidentEngine::finished(int id) const
{
    const Job &j = jobs_[(size_t)id];
    return j.therm_left + j.meas_left == 0;
}

uint32_t ResidentEngine::remaining(int id) const
{
    const Job &j = jobs_[(size_t)id];
    return j.therm_left + j.meas_left;
}

uint32_t ResidentEngine::max_remaining() const
{
    uint32_t m = 0;
    for (const Job &j : jobs_)
        if (j.live) m = std::max(m, j.therm_left + j.meas_left);
    return m;
}

int ResidentEngine::n_live() const
{
    int n = 0;
    for (const Job &j : jobs_) n += j.live ? 1 : 0;
    return n;
}

void ResidentEngine::state(int id, uint32_t &t, uint32_t &n_measured) const
{
    t = jobs_[(size_t)id].t;
    n_measured = jobs_[(size_t)id].n_measured;
}

hipError_t ResidentEngine::flush_pending()
{
    if (pending_init_.empty()) return hipSuccess;
    hipError_t e;
    // descriptors of the new lattices staged in pinned memory, two copies and
    // one admission launch for all of them (stream order puts them after any
    // launch that still reads the table; new ids are in no slot).  The staging
    // buffers are used in turn, so the host does not wait for those launches.
    const size_t n = pending_init_.size();
    const size_t idx_off = round_up(sizeof(PackedLattice) * n, 256);
    char *buf = static_cast<char *>(stage_.next(idx_off + sizeof(int32_t) * n, e));
    if (!buf) return e;
    PackedLattice *h_stage = reinterpret_cast<PackedLattice *>(buf);
    int32_t *h_idx = reinterpret_cast<int32_t *>(buf + idx_off);
    for (size_t k = 0; k < n; ++k) {
        const int id = pending_init_[k];
        const Job &job = jobs_[(size_t)id];
        h_idx[k] = id;
        PackedLattice &P = h_stage[k];
        std::memset(&P, 0, sizeof(P));
        const uint32_t W = cb_words((uint32_t)job.spec.L);
        P.spins = reinterpret_cast<uint32_t *>(arena_.base() + job.off);
        P.records = reinterpret_cast<int64_t *>(arena_.base() + job.off + round_up(2u * W * 4u, 256));
        P.k0 = (uint32_t)job.spec.seed;
        P.k1 = (uint32_t)(job.spec.seed >> 32);
        P.acc = make_accept(job.spec.coupling);
        P.t = 0;
        P.n_measured = 0;
        P.therm_left = job.spec.n_therm;
        P.meas_left = job.spec.n_meas;
        P.L = job.spec.L;
        P.n_levels = job.spec.n_levels;
        P.sweeps_per_bin = job.spec.sweeps_per_bin;
        P.max_bins = job.n_bins;
    }
    if ((e = hipMemcpyAsync(d_stage_, h_stage, sizeof(PackedLattice) * n, hipMemcpyHostToDevice, stream_)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(d_init_idx_, h_idx, sizeof(int32_t) * n, hipMemcpyHostToDevice, stream_)) != hipSuccess)
        return e;
    if ((e = stage_.used(stream_)) != hipSuccess) return e;
    if ((e = launch_packed_init(d_lats_, d_stage_, d_init_idx_, (int)n, stream_)) != hipSuccess) return e;
    pending_init_.clear();
    return hipSuccess;
}

void ResidentEngine::rebuild_slots(uint32_t chunk)
{
    struct Key {
        int L, nl;
        bool operator<(const Key &o) const { return L != o.L ? L > o.L : nl < o.nl; }
    };
    std::map<Key, std::vector<int>> groups;
    for (size_t id = 0; id < jobs_.size(); ++id) {
        const Job &j = jobs_[id];
        if (j.live && j.therm_left + j.meas_left > 0) groups[{j.spec.L, j.spec.n_levels}].push_back((int)id);
    }
    struct Pending {
        Slot s;
        uint64_t work;
        uint32_t lds;
        int threads;
    };
    std::vector<Pending> pend;
    slot_lats_.clear();
    for (auto &kv : groups) {
        const int L = kv.first.L, nl = kv.first.nl;
        const int G = group_size(L, nl, threads_for(L));
        const std::vector<int> &ids = kv.second;
        for (size_t a = 0; a < ids.size(); a += (size_t)G) {
            const int g = (int)std::min<size_t>((size_t)G, ids.size() - a);
            // positions: a power of two (so that teams tile the workgroup), empty ones are -1
            int gp = 1;
            while (gp < g) gp *= 2;
            const uint32_t ns = slot_sweeps(L, gp, chunk);
            Slot s{L, gp, nl, (int32_t)slot_lats_.size(), ns};
            uint64_t work = 0;
            for (int k = 0; k < gp; ++k) {
                if (k >= g) {
                    slot_lats_.push_back(-1);
                    continue;
                }
                const int id = ids[a + (size_t)k];
                slot_lats_.push_back(id);
                const Job &j = jobs_[(size_t)id];
                work += (uint64_t)L * (uint64_t)L * std::min<uint64_t>(ns, j.therm_left + j.meas_left);
            }
            pend.push_back({s, work, lds_bytes(L, gp, nl, threads_for(L)), threads_for(L)});
        }
    }
    // launch classes by (workgroup size, workgroups resident on the device):
    // larger workgroups first, then most resident first; the tail beyond a
    // workgroup size's share of kMaxClasses merges into its last class
    using ClassKey = std::pair<int, int>;  // (threads, resident workgroups)
    std::map<std::pair<int, uint32_t>, int> resident;  // (threads, slot LDS) -> resident workgroups
    auto key_of = [&](const Pending &p) {
        auto it = resident.find({p.threads, p.lds});
        if (it == resident.end())
            it = resident.emplace(std::make_pair(p.threads, p.lds), resident_max_active_groups(p.threads, p.lds)).first;
        return ClassKey(p.threads, it->second);
    };
    auto key_less = [](const ClassKey &a, const ClassKey &b) { return a.first != b.first ? a.first > b.first : a.second > b.second; };
    std::map<ClassKey, int, decltype(key_less)> class_of(key_less);
    for (const Pending &p : pend) class_of[key_of(p)] = 0;
    // at most kMaxClasses / (workgroup sizes in use) classes per workgroup size
    const int n_sizes = opt_.threads_large > 0 && opt_.threads_large != opt_.threads ? 2 : 1;
    const int per_size = kMaxClasses / n_sizes;
    int c = -1, last_threads = -1, in_size = 0;
    for (auto &kv : class_of) {
        if (kv.first.first != last_threads) {
            ++c;
            in_size = 1;
            last_threads = kv.first.first;
        } else if (in_size < per_size) {
            ++c;
            ++in_size;
        }
        kv.second = c;
    }
    std::vector<int> cls(pend.size());
    for (size_t i = 0; i < pend.size(); ++i) cls[i] = class_of[key_of(pend[i])];
    // within a class, longest slots first: the persistent grid pulls them in this order
    std::vector<size_t> order(pend.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return cls[a] != cls[b] ? cls[a] < cls[b] : pend[a].work > pend[b].work;
    });
    slots_.clear();
    classes_.clear();
    for (size_t i : order) {
        if (classes_.empty() || (int)classes_.size() - 1 != cls[i]) {
            LaunchClass lc;
            lc.first = (uint32_t)slots_.size();
            classes_.push_back(lc);
        }
        LaunchClass &lc = classes_.back();
        lc.count += 1;
        lc.lds = std::max(lc.lds, pend[i].lds);
        lc.threads = pend[i].threads;
        slots_.push_back(pend[i].s);
    }
}

hipError_t ResidentEngine::upload_slots()
{
    hipError_t e;
    // Growing a device table frees the old one, which a queued launch may
    // still read: only then does the host wait for the stream.
    if (slots_.size() > slot_cap_ || slot_lats_.size() > slot_lat_cap_ || slots_.size() > done_cap_)
        if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return e;
    if (slots_.size() > slot_cap_) {
        if (d_slots_) (void)hipFree(d_slots_);
        slot_cap_ = std::max<size_t>(slots_.size(), 2 * slot_cap_);
        if ((e = hipMalloc(&d_slots_, sizeof(Slot) * slot_cap_)) != hipSuccess) return e;
    }
    if (slot_lats_.size() > slot_lat_cap_) {
        if (d_slot_lats_) (void)hipFree(d_slot_lats_);
        slot_lat_cap_ = std::max<size_t>(slot_lats_.size(), 2 * slot_lat_cap_);
        if ((e = hipMalloc(&d_slot_lats_, sizeof(int32_t) * slot_lat_cap_)) != hipSuccess) return e;
    }
    if (!slots_.empty()) {
        // copies on the stream, after the launches that read the old tables,
        // from pinned staging used in turn
        const size_t lat_off = round_up(sizeof(Slot) * slots_.size(), 256);
        char *buf = static_cast<char *>(slot_stage_.next(lat_off + sizeof(int32_t) * slot_lats_.size(), e));
        if (!buf) return e;
        std::memcpy(buf, slots_.data(), sizeof(Slot) * slots_.size());
        std::memcpy(buf + lat_off, slot_lats_.data(), sizeof(int32_t) * slot_lats_.size());
        if ((e = hipMemcpyAsync(d_slots_, buf, sizeof(Slot) * slots_.size(), hipMemcpyHostToDevice, stream_)) !=
            hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(d_slot_lats_, buf + lat_off, sizeof(int32_t) * slot_lats_.size(),
                                hipMemcpyHostToDevice, stream_)) != hipSuccess)
            return e;
        if ((e = slot_stage_.used(stream_)) != hipSuccess) return e;
    }
    if (slots_.size() > done_cap_) {
        if (d_done_) (void)hipFree(d_done_);
        done_cap_ = std::max<size_t>(slots_.size(), 2 * done_cap_);
        if ((e = hipMalloc(&d_done_, sizeof(uint32_t) * done_cap_)) != hipSuccess) return e;
    }
    for (LaunchClass &lc : classes_) {
        const int resident = resident_max_active_groups(lc.threads, lc.lds);
        if (resident <= 0) {
            failed_ = "hipOccupancyMaxActiveBlocksPerMultiprocessor";
            return hipErrorInvalidConfiguration;
        }
        lc.grid = std::min<int>(resident, (int)lc.count);
        uint32_t min_sweeps = 0xFFFFFFFFu;
        for (uint32_t i = lc.first; i < lc.first + lc.count; ++i) min_sweeps = std::min(min_sweeps, slots_[i].n_sweeps);
        if (opt_.sub_chunks > 0)
            lc.n_sub = (uint32_t)opt_.sub_chunks;
        else if ((int)lc.count > cus_ && (int)lc.count < 2 * resident)
            lc.n_sub = std::max(1u, std::min(4u, min_sweeps / 16u));
        else
            lc.n_sub = 1;
        // A part waits at most for the parts of its slot queued before it,
        // i.e. for less than the slot's whole time in this launch.  Parting is
        // only a balance aid (results do not depend on it), so a class whose
        // longest slot is modelled above kPartedSlotCapNs runs unparted and no
        // wait can approach the kernel's limit.
        if (lc.n_sub > 1u) {
            uint64_t worst = 0;
            for (uint32_t i = lc.first; i < lc.first + lc.count; ++i)
                worst = std::max(worst, slot_time_ns(slots_[i].L, slots_[i].G, slots_[i].n_sweeps));
            if (worst > kPartedSlotCapNs) lc.n_sub = 1;
        }
    }
    return hipSuccess;
}

// The error word of the launches, read on stream `s` after the copies the
// caller queued there (one wait for both).  A part that gave up its wait set
// it; it is reported once and cleared, so that later launches are judged alone.
hipError_t ResidentEngine::check_parts(hipStream_t s)
{
    hipError_t e;
    if ((e = hipMemcpyAsync(h_err_, d_err_, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (*h_err_ != 0u) {
        static const uint32_t zero = 0u;
        if ((e = hipMemcpyAsync(d_err_, &zero, sizeof(zero), hipMemcpyHostToDevice, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        failed_ = "k_resident: a slot part waited for its predecessor beyond the limit";
        return hipErrorLaunchFailure;
    }
    return hipSuccess;
}

hipError_t ResidentEngine::launch_classes(KernelTimer *tm)
{
    hipError_t e;
    const int ev = tm ? tm->begin(kKindResident, stream_) : -1;
    if (classes_.size() > 1) {
        if ((e = hipEventRecord(ev_fork_, stream_)) != hipSuccess) return e;
    }
    for (size_t c = 0; c < classes_.size(); ++c) {
        const LaunchClass &lc = classes_[c];
        hipStream_t s = c == 0 ? stream_ : aux_[c - 1];
        if (c > 0 && (e = hipStreamWaitEvent(s, ev_fork_, 0)) != hipSuccess) return e;
        // one queue counter per class, 64 bytes apart
        if ((e = launch_resident(lc.threads, d_slots_ + lc.first, lc.count, lc.n_sub, d_slot_lats_, d_lats_,
                                 d_queue_ + 16u * c, d_done_ + lc.first, d_err_, lc.lds, (uint32_t)lc.grid, s)) !=
            hipSuccess) {
            failed_ = "launch_resident";
            return e;
        }
        if (c > 0) {
            if ((e = hipEventRecord(ev_join_[c], s)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(stream_, ev_join_[c], 0)) != hipSuccess) return e;
        }
    }
    if (tm) tm->end(ev, stream_);
    return hipSuccess;
}

hipError_t ResidentEngine::tick(uint32_t n, uint32_t &advanced, KernelTimer *tm)
{
    advanced = 0;
    failed_ = "";
    hipError_t e;
    if ((e = flush_pending()) != hipSuccess) {
        failed_ = "flush_pending";
        return e;
    }
    const uint32_t chunk = std::min<uint32_t>(n, (uint32_t)opt_.sweeps_per_launch);
    if (chunk == 0 || max_remaining() == 0) return hipSuccess;
    const bool timing = tuning();
    if (timing) {
        const TuneCandidate &c = tune_[tune_next_];
        opt_.threads = c.threads;
        opt_.min_team = std::min(c.min_team, c.threads);
        opt_.threads_large = c.threads_large;
        opt_.max_slot_lds = c.max_slot_lds;
        slots_dirty_ = true;
        if (!tune_a_ && (e = hipEventCreate(&tune_a_)) != hipSuccess) return e;
        if (!tune_b_ && (e = hipEventCreate(&tune_b_)) != hipSuccess) return e;
    }
    if (slots_dirty_ || chunk != slots_chunk_) {
        rebuild_slots(chunk);
        if ((e = upload_slots()) != hipSuccess) {
            if (!*failed_) failed_ = "upload_slots";
            return e;
        }
        slots_dirty_ = false;
        slots_chunk_ = chunk;
    }
    if (slots_.empty()) return hipSuccess;
    if (timing && (e = hipEventRecord(tune_a_, stream_)) != hipSuccess) return e;
    if ((e = launch_classes(tm)) != hipSuccess) {
        if (!*failed_) failed_ = "launch_classes";
        return e;
    }
    if (timing && (e = hipEventRecord(tune_b_, stream_)) != hipSuccess) return e;
    if ((e = hipEventRecord(ev_launch_[launches_ & 1u], stream_)) != hipSuccess) return e;
    ++launches_;
    // mirror the device counters
    bool any_done = false;
    double updates = 0.0;
    for (const Slot &sl : slots_) {
        for (int k = 0; k < sl.G; ++k) {
            const int id = slot_lats_[(size_t)sl.first + (size_t)k];
            if (id < 0) continue;
            Job &j = jobs_[(size_t)id];
            const uint32_t total = j.therm_left + j.meas_left;
            if (total == 0) continue;
            const uint32_t swept = std::min(total, sl.n_sweeps);
            updates += (double)sl.L * (double)sl.L * (double)swept;
            const uint32_t therm = std::min(swept, j.therm_left);
            const uint32_t meas = swept - therm;
            j.t += swept;
            j.therm_left -= therm;
            j.meas_left -= meas;
            j.n_measured += meas;
            if (j.therm_left + j.meas_left == 0) any_done = true;
        }
    }
    if (any_done) slots_dirty_ = true;
    advanced = chunk;
    if (timing) {
        float ms = 0.f;
        if ((e = hipEventSynchronize(tune_b_)) != hipSuccess) return e;
        if ((e = hipEventElapsedTime(&ms, tune_a_, tune_b_)) != hipSuccess) return e;
        tune_[tune_next_].rate = ms > 0.f ? updates / (double)ms : 0.0;
        if (++tune_next_ == tune_.size()) {
            // entry 0 only warmed up; entry 1 (the configuration in use) is
            // replaced only by a candidate faster by more than one launch's
            // timing noise (up to 3 % between launches of the batched
            // workload), so near-equal candidates do not flip the choice from
            // run to run
            size_t best = 2;
            for (size_t i = 3; i < tune_.size(); ++i)
                if (tune_[i].rate > tune_[best].rate) best = i;
            if (tune_[best].rate <= kTuneMargin * tune_[1].rate) best = 1;
            opt_.threads = tune_[best].threads;
            opt_.min_team = std::min(tune_[best].min_team, tune_[best].threads);
            opt_.threads_large = tune_[best].threads_large;
            opt_.max_slot_lds = tune_[best].max_slot_lds;
            slots_dirty_ = true;
        }
    }
    return hipSuccess;
}

hipError_t ResidentEngine::read_spins(int id, int8_t *host)
{
    hipError_t e;
    if ((e = flush_pending()) != hipSuccess) return e;
    const Job &job = jobs_[(size_t)id];
    const uint32_t L = (uint32_t)job.spec.L, W = cb_words(L), half = L / 2;
    std::vector<uint32_t> words(2u * W);
    if ((e = hipMemcpyAsync(words.data(), arena_.base() + job.off, words.size() * 4u, hipMemcpyDeviceToHost,
                            stream_)) != hipSuccess)
        return e;
    if ((e = check_parts(stream_)) != hipSuccess) return e;
    for (uint32_t c = 0; c < 2u; ++c)
        for (uint32_t w = 0; w < W; ++w)
            for (uint32_t b = 0; b < 32u; ++b) {
                const uint32_t mc = 32u * w + b;
                const uint32_t y = mc / half, k = mc % half;
                const uint32_t x = 2u * k + ((y + c) & 1u);
                host[y * L + x] = ((words[c * W + w] >> b) & 1u) ? -1 : 1;
            }
    return hipSuccess;
}

hipError_t ResidentEngine::copy_records(hipStream_t s, const std::vector<int> &ids,
                                        const std::vector<int64_t *> &hosts, const std::vector<int64_t> &ns)
{
    hipError_t e;
    size_t total = 0;
    for (int64_t n : ns) total += (size_t)n;
    if (total > rec_cap_) {
        if (h_rec_) (void)hipHostFree(h_rec_);
        h_rec_ = nullptr;
        rec_cap_ = 0;
        const size_t cap = std::max<size_t>(total, 1u << 16);
        if ((e = hipHostMalloc(reinterpret_cast<void **>(&h_rec_), cap * 8u, hipHostMallocDefault)) != hipSuccess)
            return e;
        rec_cap_ = cap;
    }
    // every copy queued into pinned memory, one wait for all of them (and the
    // error word)
    size_t off = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
        const Job &job = jobs_[(size_t)ids[k]];
        const uint32_t W = cb_words((uint32_t)job.spec.L);
        const char *rec = arena_.base() + job.off + round_up(2u * W * 4u, 256);
        if ((e = hipMemcpyAsync(h_rec_ + off, rec, (size_t)ns[k] * 8u, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return e;
        off += (size_t)ns[k];
    }
    if ((e = check_parts(s)) != hipSuccess) return e;
    off = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
        std::memcpy(hosts[k], h_rec_ + off, (size_t)ns[k] * 8u);
        off += (size_t)ns[k];
    }
    return hipSuccess;
}

hipError_t ResidentEngine::read_records_many(const std::vector<int> &ids, const std::vector<int64_t *> &hosts,
                                             const std::vector<int64_t> &ns)
{
    hipError_t e;
    if ((e = flush_pending()) != hipSuccess) return e;
    return copy_records(stream_, ids, hosts, ns);
}

hipError_t ResidentEngine::wait_launch(uint64_t k)
{
    if (k >= launches_ || launches_ - k > 2u) return hipErrorInvalidValue;
    return hipEventSynchronize(ev_launch_[k & 1u]);
}

hipError_t ResidentEngine::read_records_after(uint64_t k, const std::vector<int> &ids,
                                              const std::vector<int64_t *> &hosts, const std::vector<int64_t> &ns)
{
    hipError_t e;
    // launch k (the event behind it; later launches keep running) ...
    if ((e = wait_launch(k)) != hipSuccess) return e;
    // ... then the copies on the readback stream, which no launch waits behind
    return copy_records(rb_, ids, hosts, ns);
}

hipError_t ResidentEngine::read_records(int id, int64_t *host, int64_t n)
{
    hipError_t e;
    if ((e = flush_pending()) != hipSuccess) return e;
    const Job &job = jobs_[(size_t)id];
    const uint32_t W = cb_words((uint32_t)job.spec.L);
    const char *rec = arena_.base() + job.off + round_up(2u * W * 4u, 256);
    if ((e = hipMemcpyAsync(host, rec, (size_t)n * 8u, hipMemcpyDeviceToHost, stream_)) != hipSuccess) return e;
    return check_parts(stream_);
}

}  // namespace mcrg
