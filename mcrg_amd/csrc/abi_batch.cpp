// C-ABI: batches of independent chains.  Metropolis lattices of 8 <= L <= 512
// run on the LDS-resident schedule (bit-packed, one persistent launch per
// chunk); the others (L = 2, 4, L >= 1024, and Swendsen-Wang jobs) run on the
// HBM schedule (int8), grouped by
// shape so that each group advances through shared launches, on a second
// stream that overlaps the resident launches.
#include <cstdlib>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "hbm_group.hpp"
#include "runtime.hpp"

using namespace mcrg;

namespace {
struct JobRef {
    int group;  // -1: resident engine
    int index;  // engine id, or member of the group
};

// Metropolis jobs of kResidentMinL <= L <= kResidentMaxL run on the resident
// engine, the rest in int8 HBM-schedule groups.  MCRG_BATCH_RESIDENT=0 in the
// environment sends every job to the HBM schedule (same results), to compare
// the two schedules.
bool resident_job(const JobSpec &s)
{
    static const bool allowed = [] {
        const char *v = std::getenv("MCRG_BATCH_RESIDENT");
        return !(v && std::atoi(v) == 0);
    }();
    return allowed && s.sampler == MCRG_SAMPLER_METROPOLIS && s.L >= kResidentMinL && s.L <= kResidentMaxL;
}
}  // namespace

struct mcrg_batch_s {
    mcrg_ctx ctx = nullptr;
    hipStream_t stream = nullptr;      // resident launches
    hipStream_t hbm_stream = nullptr;  // HBM-schedule groups
    ResidentEngine engine;
    int n_resident = 0;
    uint32_t chunk = 250;
    std::vector<std::unique_ptr<HbmGroup>> groups;
    std::vector<JobRef> refs;  // job index -> where it runs

    ~mcrg_batch_s()
    {
        if (hbm_stream) (void)hipStreamSynchronize(hbm_stream);
        groups.clear();
    }
    uint32_t max_remaining() const
    {
        uint32_t m = n_resident ? engine.max_remaining() : 0u;
        for (const auto &g : groups) m = std::max(m, g->remaining());
        return m;
    }
};

static mcrg_status batch_job(mcrg_batch b, int32_t job)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_REQUIRE(job >= 0 && job < (int32_t)b->refs.size(), "job index out of range");
    return MCRG_OK;
}

extern "C" {

mcrg_status mcrg_batch_create(mcrg_ctx ctx, const mcrg_job_desc *jobs, int32_t n_jobs,
                              const mcrg_batch_opts *opts, mcrg_batch *out)
{
    MCRG_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
    MCRG_REQUIRE(n_jobs >= 1 && jobs != nullptr, "need at least one job");
    EngineOptions eo;
    if (opts) {
        if (opts->threads) eo.threads = opts->threads;
        if (opts->sweeps_per_launch) eo.sweeps_per_launch = opts->sweeps_per_launch;
        if (opts->max_group) eo.max_group = opts->max_group;
        MCRG_REQUIRE(opts->max_slot_lds >= 0, "max_slot_lds must be >= 0");
        eo.max_slot_lds = (uint32_t)opts->max_slot_lds;
        if (opts->min_team) eo.min_team = opts->min_team;
        eo.max_team = opts->max_team;
        eo.threads_large = opts->threads_large;
        eo.sub_chunks = opts->sub_chunks;
    }
    MCRG_REQUIRE(eo.sub_chunks >= 0 && eo.sub_chunks <= 64, "sub_chunks must be in [0, 64]");
    MCRG_REQUIRE(eo.threads_large == 0 || eo.threads_large == 64 || eo.threads_large == 128 ||
                     eo.threads_large == 256 || eo.threads_large == 512 || eo.threads_large == 1024,
                 "threads_large must be 0, 64, 128, 256, 512 or 1024");
    MCRG_REQUIRE(eo.max_team == 0 || (eo.max_team >= eo.min_team && eo.max_team <= eo.threads &&
                                      (eo.max_team & (eo.max_team - 1)) == 0),
                 "max_team must be 0 or a power of two in [min_team, threads]");
    MCRG_REQUIRE(eo.min_team >= 1 && eo.min_team <= eo.threads && (eo.min_team & (eo.min_team - 1)) == 0,
                 "min_team must be a power of two in [1, threads]");
    MCRG_REQUIRE(eo.threads == 64 || eo.threads == 128 || eo.threads == 256 || eo.threads == 512,
                 "threads must be 64, 128, 256 or 512");
    MCRG_REQUIRE(eo.sweeps_per_launch >= 1, "sweeps_per_launch must be >= 1");
    MCRG_REQUIRE(eo.max_group >= 1 && eo.max_group <= 4096, "max_group must be in [1, 4096]");
    std::vector<JobSpec> specs((size_t)n_jobs);
    size_t bytes = 0;
    int n_resident = 0, n_large = 0;
    for (int32_t i = 0; i < n_jobs; ++i) {
        const mcrg_job_desc &d = jobs[i];
        MCRG_REQUIRE(d.n_therm >= 0 && d.n_meas >= 0, "sweep counts must be >= 0");
        JobSpec s;
        s.L = d.L;
        s.coupling = d.coupling;
        s.seed = d.seed;
        s.n_therm = (uint32_t)d.n_therm;
        s.n_meas = (uint32_t)d.n_meas;
        s.n_levels = d.n_levels;
        s.sweeps_per_bin = d.sweeps_per_bin;
        s.sampler = d.sampler;
        const std::string why = resident_job(s) ? ResidentEngine::check(s) : HbmGroup::check(s);
        if (!why.empty()) {
            set_error("job " + std::to_string(i) + ": " + why);
            return MCRG_ERR_INVALID_ARG;
        }
        specs[(size_t)i] = s;
        if (resident_job(s)) {
            bytes += ResidentEngine::job_bytes(s);
            ++n_resident;
            if (s.L >= eo.large_min_L) ++n_large;
        }
    }
    MCRG_HIP(hipSetDevice(ctx->device));
    if (eo.threads_large == 0 && n_large > 0) {
        // L = 512 lattices: 1024-lane workgroups while there are no more of
        // them than compute units (one per CU; each sweeps 1.2x faster than
        // on 512 lanes), 512-lane ones beyond (more of them fit a CU).  With
        // 256 L = 512 lattices: 1781 / 1480 / 1002 updates/ns on 1024 / 512 /
        // 256 lanes; with 1024: 1855 / 2209 / 1718.
        int cus = 0;
        MCRG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        eo.threads_large = n_large <= cus ? 1024 : 512;
    }
    mcrg_batch b = new (std::nothrow) mcrg_batch_s;
    if (!b) {
        set_error("out of host memory");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    b->ctx = ctx;
    b->n_resident = n_resident;
    b->chunk = (uint32_t)eo.sweeps_per_launch;
    auto fail = [&](hipError_t e, const char *what) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        hipStream_t s1 = b->stream, s2 = b->hbm_stream;
        delete b;
        if (s1) (void)hipStreamDestroy(s1);
        if (s2) (void)hipStreamDestroy(s2);
        return e == hipErrorOutOfMemory ? MCRG_ERR_OUT_OF_BUDGET : MCRG_ERR_DEVICE;
    };
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->hbm_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = b->engine.init(ctx->device, std::max<size_t>(bytes, 256), std::max(n_resident, 1), eo,
                                            b->stream);
    if (e != hipSuccess) return fail(e, "batch allocation");
    b->refs.resize((size_t)n_jobs);
    // HBM-schedule jobs: one group per shape, members in job order
    std::map<std::tuple<int, int, uint32_t, uint32_t, int, int>, int> group_of;
    std::vector<std::vector<int>> members;
    for (int32_t i = 0; i < n_jobs; ++i) {
        const JobSpec &s = specs[(size_t)i];
        if (resident_job(s)) {
            const int id = b->engine.add(s);
            if (id < 0) {
                set_error("batch arena exhausted");
                delete b;
                return MCRG_ERR_OUT_OF_BUDGET;
            }
            b->refs[(size_t)i] = {-1, id};
            continue;
        }
        const auto key = std::make_tuple(s.L, s.n_levels, s.n_therm, s.n_meas, s.sweeps_per_bin, s.sampler);
        auto it = group_of.find(key);
        if (it == group_of.end()) {
            it = group_of.emplace(key, (int)members.size()).first;
            members.emplace_back();
        }
        b->refs[(size_t)i] = {it->second, (int)members[(size_t)it->second].size()};
        members[(size_t)it->second].push_back(i);
    }
    for (const std::vector<int> &m : members) {
        std::vector<JobSpec> ms;
        for (int i : m) ms.push_back(specs[(size_t)i]);
        b->groups.emplace_back(new HbmGroup());
        if ((e = b->groups.back()->init(ms, b->hbm_stream)) != hipSuccess) return fail(e, "batch HBM group");
    }
    if ((e = b->engine.flush_pending()) != hipSuccess) return fail(e, "batch init");
    if (opts && opts->autotune && n_resident > 0) b->engine.enable_autotune();
    *out = b;
    return MCRG_OK;
}

mcrg_status mcrg_batch_destroy(mcrg_batch b)
{
    if (!b) return MCRG_OK;
    (void)hipSetDevice(b->ctx->device);
    hipStream_t s1 = b->stream, s2 = b->hbm_stream;
    delete b;  // the engine and the groups synchronise their streams before freeing
    if (s1) (void)hipStreamDestroy(s1);
    if (s2) (void)hipStreamDestroy(s2);
    return MCRG_OK;
}

mcrg_status mcrg_batch_run(mcrg_batch b, int32_t max_sweeps, int32_t *remaining)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_REQUIRE(max_sweeps >= 0, "max_sweeps must be >= 0");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    uint32_t left = (uint32_t)max_sweeps;
    KernelTimer *tm = b->ctx->timer.enabled ? &b->ctx->timer : nullptr;
    while (left > 0 && b->max_remaining() > 0) {
        // HBM groups first: their launches queue on the second stream while the
        // host prepares and launches the resident chunk
        const uint32_t step = std::min(left, b->chunk);
        for (auto &g : b->groups) {
            if (g->remaining() == 0) continue;
            const hipError_t e = g->advance(step, b->hbm_stream, tm);
            if (e != hipSuccess) {
                set_error(std::string("batch HBM group: ") + hipGetErrorString(e));
                return MCRG_ERR_DEVICE;
            }
        }
        if (b->n_resident) {
            uint32_t adv = 0;
            const hipError_t e = b->engine.tick(step, adv, tm);
            if (e != hipSuccess) {
                set_error(std::string("batch tick (") + b->engine.failed_call() + "): " + hipGetErrorString(e));
                return MCRG_ERR_DEVICE;
            }
        }
        left -= step;
    }
    if (remaining) *remaining = (int32_t)b->max_remaining();
    return MCRG_OK;
}

mcrg_status mcrg_batch_sync(mcrg_batch b)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    MCRG_HIP(hipStreamSynchronize(b->stream));
    MCRG_HIP(hipStreamSynchronize(b->hbm_stream));
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_records(mcrg_batch b, int32_t job, int64_t *host, int64_t n)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    MCRG_REQUIRE(host != nullptr, "host is null");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    const JobRef r = b->refs[(size_t)job];
    if (r.group < 0) {
        const int64_t want = (int64_t)b->engine.n_bins(r.index) * record_size((uint32_t)b->engine.spec(r.index).n_levels);
        MCRG_REQUIRE(n == want, "n must be n_bins * mcrg_record_size(n_levels)");
        MCRG_HIP(b->engine.read_records(r.index, host, n));
    } else {
        HbmGroup &g = *b->groups[(size_t)r.group];
        const int64_t want = (int64_t)g.n_bins() * record_size((uint32_t)g.spec().n_levels);
        MCRG_REQUIRE(n == want, "n must be n_bins * mcrg_record_size(n_levels)");
        MCRG_HIP(g.read_records(r.index, host, n, b->hbm_stream));
    }
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_spins(mcrg_batch b, int32_t job, int8_t *host, int64_t n)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    MCRG_REQUIRE(host != nullptr, "host is null");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    const JobRef r = b->refs[(size_t)job];
    if (r.group < 0) {
        const int L = b->engine.spec(r.index).L;
        MCRG_REQUIRE(n == (int64_t)L * L, "n must be L*L");
        MCRG_HIP(b->engine.read_spins(r.index, host));
    } else {
        HbmGroup &g = *b->groups[(size_t)r.group];
        MCRG_REQUIRE(n == (int64_t)g.spec().L * g.spec().L, "n must be L*L");
        MCRG_HIP(g.read_spins(r.index, host, b->hbm_stream));
    }
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_opts(mcrg_batch b, mcrg_batch_opts *out)
{
    MCRG_REQUIRE(b != nullptr && out != nullptr, "null argument");
    const EngineOptions &o = b->engine.options();
    out->threads = o.threads;
    out->sweeps_per_launch = o.sweeps_per_launch;
    out->max_group = o.max_group;
    out->max_slot_lds = (int32_t)o.max_slot_lds;
    out->min_team = o.min_team;
    out->max_team = o.max_team;
    out->threads_large = o.threads_large;
    out->autotune = b->engine.tuning() ? 1 : 0;
    out->sub_chunks = o.sub_chunks;
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_state(mcrg_batch b, int32_t job, uint32_t *sweep, uint32_t *n_measured)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    uint32_t t = 0, q = 0;
    const JobRef r = b->refs[(size_t)job];
    if (r.group < 0)
        b->engine.state(r.index, t, q);
    else
        b->groups[(size_t)r.group]->state(t, q);
    if (sweep) *sweep = t;
    if (n_measured) *n_measured = q;
    return MCRG_OK;
}

}  // extern "C"
