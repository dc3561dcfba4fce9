// C-ABI: batches of independent chains on the LDS-resident schedule.
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "runtime.hpp"

using namespace mcrg;

struct mcrg_batch_s {
    mcrg_ctx ctx = nullptr;
    hipStream_t stream = nullptr;
    ResidentEngine engine;
    std::vector<int> ids;  // job index -> engine id
};

static mcrg_status batch_job(mcrg_batch b, int32_t job)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_REQUIRE(job >= 0 && job < (int32_t)b->ids.size(), "job index out of range");
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
    }
    MCRG_REQUIRE(eo.threads == 64 || eo.threads == 128 || eo.threads == 256 || eo.threads == 512,
                 "threads must be 64, 128, 256 or 512");
    MCRG_REQUIRE(eo.sweeps_per_launch >= 1, "sweeps_per_launch must be >= 1");
    MCRG_REQUIRE(eo.max_group >= 1 && eo.max_group <= 4096, "max_group must be in [1, 4096]");
    std::vector<JobSpec> specs((size_t)n_jobs);
    size_t bytes = 0;
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
        const std::string why = ResidentEngine::check(s);
        if (!why.empty()) {
            set_error("job " + std::to_string(i) + ": " + why);
            return MCRG_ERR_INVALID_ARG;
        }
        specs[(size_t)i] = s;
        bytes += ResidentEngine::job_bytes(s);
    }
    MCRG_HIP(hipSetDevice(ctx->device));
    mcrg_batch b = new (std::nothrow) mcrg_batch_s;
    if (!b) {
        set_error("out of host memory");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    b->ctx = ctx;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = b->engine.init(ctx->device, bytes, n_jobs, eo, b->stream);
    if (e != hipSuccess) {
        set_error(std::string("batch allocation: ") + hipGetErrorString(e));
        delete b;
        return e == hipErrorOutOfMemory ? MCRG_ERR_OUT_OF_BUDGET : MCRG_ERR_DEVICE;
    }
    b->ids.resize((size_t)n_jobs);
    for (int32_t i = 0; i < n_jobs; ++i) {
        const int id = b->engine.add(specs[(size_t)i]);
        if (id < 0) {
            set_error("batch arena exhausted");
            delete b;
            return MCRG_ERR_OUT_OF_BUDGET;
        }
        b->ids[(size_t)i] = id;
    }
    e = b->engine.flush_pending();
    if (e != hipSuccess) {
        set_error(std::string("batch init: ") + hipGetErrorString(e));
        delete b;
        return MCRG_ERR_DEVICE;
    }
    *out = b;
    return MCRG_OK;
}

mcrg_status mcrg_batch_destroy(mcrg_batch b)
{
    if (!b) return MCRG_OK;
    (void)hipSetDevice(b->ctx->device);
    hipStream_t s = b->stream;
    delete b;  // the engine synchronises its stream before freeing
    if (s) (void)hipStreamDestroy(s);
    return MCRG_OK;
}

mcrg_status mcrg_batch_run(mcrg_batch b, int32_t max_sweeps, int32_t *remaining)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_REQUIRE(max_sweeps >= 0, "max_sweeps must be >= 0");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    uint32_t left = (uint32_t)max_sweeps;
    KernelTimer *tm = b->ctx->timer.enabled ? &b->ctx->timer : nullptr;
    while (left > 0) {
        uint32_t adv = 0;
        const hipError_t e = b->engine.tick(left, adv, tm);
        if (e != hipSuccess) {
            set_error(std::string("batch tick (") + b->engine.failed_call() + "): " + hipGetErrorString(e));
            return MCRG_ERR_DEVICE;
        }
        if (adv == 0) break;
        left -= adv;
    }
    if (remaining) *remaining = (int32_t)b->engine.max_remaining();
    return MCRG_OK;
}

mcrg_status mcrg_batch_sync(mcrg_batch b)
{
    MCRG_REQUIRE(b != nullptr, "batch is null");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    MCRG_HIP(hipStreamSynchronize(b->stream));
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_records(mcrg_batch b, int32_t job, int64_t *host, int64_t n)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    MCRG_REQUIRE(host != nullptr, "host is null");
    const int id = b->ids[(size_t)job];
    const int64_t want = (int64_t)b->engine.n_bins(id) * record_size((uint32_t)b->engine.spec(id).n_levels);
    MCRG_REQUIRE(n == want, "n must be n_bins * mcrg_record_size(n_levels)");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    MCRG_HIP(b->engine.read_records(id, host, n));
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_spins(mcrg_batch b, int32_t job, int8_t *host, int64_t n)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    MCRG_REQUIRE(host != nullptr, "host is null");
    const int id = b->ids[(size_t)job];
    const int L = b->engine.spec(id).L;
    MCRG_REQUIRE(n == (int64_t)L * L, "n must be L*L");
    MCRG_HIP(hipSetDevice(b->ctx->device));
    MCRG_HIP(b->engine.read_spins(id, host));
    return MCRG_OK;
}

mcrg_status mcrg_batch_get_state(mcrg_batch b, int32_t job, uint32_t *sweep, uint32_t *n_measured)
{
    mcrg_status st = batch_job(b, job);
    if (st != MCRG_OK) return st;
    uint32_t t = 0, q = 0;
    b->engine.state(b->ids[(size_t)job], t, q);
    if (sweep) *sweep = t;
    if (n_measured) *n_measured = q;
    return MCRG_OK;
}

}  // extern "C"
