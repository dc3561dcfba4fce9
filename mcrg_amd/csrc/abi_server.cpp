// C-ABI: continuous-batching job server over the LDS-resident engine.
//
// One scheduler (a thread of its own after mcrg_server_start, or the caller of
// mcrg_server_step) owns the engine and the device stream.  Each iteration:
//   admit   queued jobs in arrival order while the device-memory and LDS
//           budgets have room
//   launch  one time slice for every resident job that still has sweeps
//   retire  the jobs the previous launch finished: wait for that launch
//           alone, records to host, memory freed
// so the next launch is already queued while a launch's jobs are retired and
// new ones admitted, and the device does not wait for the host.
// Submitting, polling and reading results only touch the host job table.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "runtime.hpp"

using namespace mcrg;

namespace {
struct ServerJob {
    JobSpec spec;
    int state = MCRG_JOB_QUEUED;
    int engine_id = -1;
    int n_bins = 0;
    uint32_t lds = 0;  // LDS share while resident
    bool cancel = false;
    bool retiring = false;  // finished by the last launch, records not read yet
    double t_submit = 0.0, t_start = 0.0, t_finish = 0.0;
    std::vector<int64_t> records;
};
}  // namespace

struct mcrg_server_s {
    mcrg_ctx ctx = nullptr;
    hipStream_t stream = nullptr;
    ResidentEngine engine;
    EngineOptions eo;
    std::mutex mu;
    std::condition_variable cv;
    std::unordered_map<uint64_t, ServerJob> jobs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;
    std::deque<uint64_t> completed;
    uint64_t next_id = 0;
    int64_t lds_budget = 0, lds_in_use = 0;
    bool adaptive_large = false;  // large-lattice workgroup size follows the queue
    std::thread th;
    std::atomic<bool> stop{false};
    bool started = false;
    std::mutex step_mu;  // one scheduling iteration at a time
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    mcrg_server_stats st{};
    std::string fatal;
    // jobs the launch `retire_launch` finished, retired after the next launch
    std::vector<uint64_t> retire;
    uint64_t retire_launch = 0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static mcrg_status server_step(mcrg_server s, int32_t *active)
{
    std::lock_guard<std::mutex> step_lock(s->step_mu);
    // ---- cancellations of resident jobs and admissions (host tables under the lock)
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (size_t i = 0; i < s->running.size();) {
            ServerJob &j = s->jobs.at(s->running[i]);  // running jobs are never erased
            if (j.cancel) {
                s->engine.remove(j.engine_id);
                s->lds_in_use -= j.lds;
                j.state = MCRG_JOB_CANCELLED;
                j.t_finish = s->now();
                s->st.cancelled++;
                s->running[i] = s->running.back();
                s->running.pop_back();
            } else {
                ++i;
            }
        }
        while (!s->queue.empty()) {
            const uint64_t id = s->queue.front();
            ServerJob &j = s->jobs.at(id);  // queued jobs are never erased
            if (j.cancel) {
                j.state = MCRG_JOB_CANCELLED;
                s->st.cancelled++;
                s->queue.pop_front();
                continue;
            }
            const uint32_t lds = s->engine.lds_share(j.spec.L, j.spec.n_levels);
            if (s->lds_in_use + (int64_t)lds > s->lds_budget) break;  // wait for retirements
            const int eid = s->engine.add(j.spec);
            if (eid < 0) break;  // device memory or the job table is full
            j.engine_id = eid;
            j.lds = lds;
            s->lds_in_use += lds;
            j.n_bins = s->engine.n_bins(eid);
            j.state = MCRG_JOB_RUNNING;
            j.t_start = s->now();
            s->st.admitted++;
            s->running.push_back(id);
            s->queue.pop_front();
        }
        // Latency mode while every job is admitted: the long L >= 256 jobs get
        // 1024-lane workgroups (each lattice sweeps 1.2x faster).  Once jobs
        // wait for room the device is saturated and 512-lane workgroups give
        // more aggregate updates (2237 against 1864 updates/ns with 512
        // resident L = 512 lattices).
        if (s->adaptive_large) s->engine.set_threads_large(s->queue.empty() ? 1024 : 512);
        s->st.queued = (int64_t)s->queue.size();
        s->st.running = (int64_t)s->running.size();
        s->st.hbm_in_use = (int64_t)s->engine.arena_in_use();
        s->st.lds_in_use = s->lds_in_use;
    }
    if (s->running.empty()) {
        if (active) *active = 0;
        return MCRG_OK;
    }
    // ---- one chunk for every resident job that still has sweeps (the jobs
    //      waiting for retirement have none); not waited for here
    const double t_launch = s->now();
    uint32_t adv = 0;
    KernelTimer *tm = s->ctx->timer.enabled ? &s->ctx->timer : nullptr;
    const uint64_t before = s->engine.launches();
    hipError_t e = s->engine.tick((uint32_t)s->eo.sweeps_per_launch, adv, tm);
    (void)adv;
    if (e != hipSuccess) {
        set_error(std::string("server launch (") + s->engine.failed_call() + "): " + hipGetErrorString(e));
        return MCRG_ERR_DEVICE;
    }
    const bool launched = s->engine.launches() != before;
    // ---- retire the jobs the previous launch finished.  The job table is
    //      shared with client threads (submit inserts, release erases): it is
    //      only touched under `mu`; device reads happen outside the lock.
    struct Done {
        uint64_t id;
        int engine_id;
        std::vector<int64_t> rec;
        double updates;
    };
    std::vector<Done> done;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (uint64_t id : s->retire) {
            auto it = s->jobs.find(id);
            if (it == s->jobs.end() || it->second.state != MCRG_JOB_RUNNING) continue;  // cancelled meanwhile
            const ServerJob &j = it->second;
            Done d;
            d.id = id;
            d.engine_id = j.engine_id;
            d.rec.resize((size_t)j.n_bins * record_size((uint32_t)j.spec.n_levels));
            d.updates = (double)j.spec.L * j.spec.L * (double)(j.spec.n_therm + j.spec.n_meas);
            done.push_back(std::move(d));
        }
        s->retire.clear();
    }
    const double t_wait = s->now();
    // at most two launches in flight: the one just queued and the one before,
    // which this iteration waits for (its finished jobs are read back now)
    if (launched && s->engine.launches() >= 2u && (e = s->engine.wait_launch(s->engine.launches() - 2u)) != hipSuccess) {
        set_error(std::string("server launch: ") + hipGetErrorString(e));
        return MCRG_ERR_DEVICE;
    }
    if (!done.empty()) {
        std::vector<int> ids;
        std::vector<int64_t *> dst;
        std::vector<int64_t> ns;
        for (Done &d : done) {
            ids.push_back(d.engine_id);
            dst.push_back(d.rec.data());
            ns.push_back((int64_t)d.rec.size());
        }
        e = s->engine.read_records_after(s->retire_launch, ids, dst, ns);
        if (e != hipSuccess) {
            set_error(std::string("server readback (") + s->engine.failed_call() + "): " + hipGetErrorString(e));
            return MCRG_ERR_DEVICE;
        }
        for (Done &d : done) s->engine.remove(d.engine_id);
    }
    const double t_done = s->now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = s->now();
        for (Done &d : done) {
            auto it = s->jobs.find(d.id);
            if (it == s->jobs.end() || it->second.state != MCRG_JOB_RUNNING) continue;
            ServerJob &j = it->second;
            s->lds_in_use -= j.lds;
            j.records.swap(d.rec);
            j.state = MCRG_JOB_DONE;
            j.t_finish = t;
            s->completed.push_back(d.id);
            s->st.completed++;
            s->st.site_updates += d.updates;
        }
        if (!done.empty()) {
            size_t k = 0;
            for (uint64_t id : s->running) {
                auto it = s->jobs.find(id);
                if (it != s->jobs.end() && it->second.state == MCRG_JOB_RUNNING) s->running[k++] = id;
            }
            s->running.resize(k);
        }
        // the jobs this launch finishes (the host mirrors the device counters
        // exactly), retired in the next iteration once it has run
        if (launched) {
            for (uint64_t id : s->running) {
                ServerJob &j = s->jobs.at(id);  // running jobs are never erased
                if (!j.retiring && s->engine.finished(j.engine_id)) {
                    j.retiring = true;
                    s->retire.push_back(id);
                }
            }
            s->retire_launch = s->engine.launches() - 1u;
            s->st.launches++;
        }
        s->st.busy_s += s->now() - t_launch;
        s->st.device_wait_s += t_done - t_wait;
        s->st.running = (int64_t)s->running.size();
        s->st.hbm_in_use = (int64_t)s->engine.arena_in_use();
        s->st.lds_in_use = s->lds_in_use;
        if (active) *active = (int32_t)s->running.size();
    }
    s->cv.notify_all();
    return MCRG_OK;
}

static void server_loop(mcrg_server s)
{
    (void)hipSetDevice(s->ctx->device);
    while (!s->stop.load()) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->queue.empty() && s->running.empty()) {
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
        }
        int32_t active = 0;
        if (server_step(s, &active) != MCRG_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->fatal = mcrg_last_error();
            break;
        }
    }
}

extern "C" {

mcrg_status mcrg_server_create(mcrg_ctx ctx, const mcrg_server_opts *opts, mcrg_server *out)
{
    MCRG_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
    int64_t budget = int64_t(4) << 30, lds_budget = 0;
    int max_jobs = 65536;
    EngineOptions eo;
    eo.sweeps_per_launch = 65536;
    eo.quantum_ns = 4000000;  // 4 ms of work per workgroup and launch
    // L = 256 and 512 jobs are the longest and set the tail latency: one
    // L = 512 lattice sweeps 1.2x faster on 1024 lanes than on 512 (and 1.5x
    // faster on 512 than on 256).  Unless the caller fixes it, the size
    // follows the queue (server_step).
    eo.threads_large = 1024;
    eo.large_min_L = 256;
    bool adaptive = true;
    if (opts) {
        if (opts->hbm_budget_bytes) budget = opts->hbm_budget_bytes;
        if (opts->lds_budget_bytes) lds_budget = opts->lds_budget_bytes;
        if (opts->launch_quantum) eo.quantum_ns = (uint64_t)opts->launch_quantum * 1000u;
        if (opts->max_jobs) max_jobs = opts->max_jobs;
        if (opts->threads) eo.threads = opts->threads;
        if (opts->sweeps_per_launch) eo.sweeps_per_launch = opts->sweeps_per_launch;
        MCRG_REQUIRE(opts->launch_quantum >= 0, "launch_quantum must be >= 0");
        MCRG_REQUIRE(opts->lds_budget_bytes >= 0, "lds_budget_bytes must be >= 0");
        if (opts->threads_large) {
            eo.threads_large = opts->threads_large;
            adaptive = false;
        }
    }
    MCRG_REQUIRE(budget >= 4096, "hbm_budget_bytes must be >= 4096");
    MCRG_REQUIRE(max_jobs >= 1, "max_jobs must be >= 1");
    MCRG_REQUIRE(eo.threads == 64 || eo.threads == 128 || eo.threads == 256 || eo.threads == 512,
                 "threads must be 64, 128, 256 or 512");
    MCRG_REQUIRE(eo.sweeps_per_launch >= 1, "sweeps_per_launch must be >= 1");
    MCRG_REQUIRE(eo.threads_large == 64 || eo.threads_large == 128 || eo.threads_large == 256 ||
                     eo.threads_large == 512 || eo.threads_large == 1024,
                 "threads_large must be 64, 128, 256, 512 or 1024");
    MCRG_HIP(hipSetDevice(ctx->device));
    if (lds_budget == 0) {
        int cus = 0, lds_cu = 0;
        MCRG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        MCRG_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx->device));
        lds_budget = 2 * (int64_t)cus * (int64_t)lds_cu;
    }
    mcrg_server s = new (std::nothrow) mcrg_server_s;
    if (!s) {
        set_error("out of host memory");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    s->ctx = ctx;
    s->eo = eo;
    s->lds_budget = lds_budget;
    s->adaptive_large = adaptive;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = s->engine.init(ctx->device, (size_t)budget, max_jobs, eo, s->stream);
    if (e != hipSuccess) {
        set_error(std::string("server allocation: ") + hipGetErrorString(e));
        delete s;
        return e == hipErrorOutOfMemory ? MCRG_ERR_OUT_OF_BUDGET : MCRG_ERR_DEVICE;
    }
    *out = s;
    return MCRG_OK;
}

mcrg_status mcrg_server_destroy(mcrg_server s)
{
    if (!s) return MCRG_OK;
    mcrg_server_stop(s);
    (void)hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
    return MCRG_OK;
}

mcrg_status mcrg_server_submit(mcrg_server s, const mcrg_job_desc *job, uint64_t *job_id)
{
    MCRG_REQUIRE(s != nullptr && job != nullptr && job_id != nullptr, "null argument");
    MCRG_REQUIRE(job->n_therm >= 0 && job->n_meas >= 0, "sweep counts must be >= 0");
    JobSpec spec;
    spec.L = job->L;
    spec.coupling = job->coupling;
    spec.seed = job->seed;
    spec.n_therm = (uint32_t)job->n_therm;
    spec.n_meas = (uint32_t)job->n_meas;
    spec.n_levels = job->n_levels;
    spec.sweeps_per_bin = job->sweeps_per_bin;
    spec.sampler = job->sampler;
    const std::string why = ResidentEngine::check(spec);
    if (!why.empty()) {
        set_error(why);
        return MCRG_ERR_INVALID_ARG;
    }
    if (ResidentEngine::job_bytes(spec) > s->engine.arena_capacity()) {
        set_error("job needs more device memory than the whole budget");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);  // the engine's options change in server_step
        // the job has to fit the budget on every workgroup size the server may use
        uint32_t share = s->engine.lds_share(spec.L, spec.n_levels);
        if (s->adaptive_large && spec.L >= s->eo.large_min_L)
            share = std::max(s->engine.lds_share(spec.L, spec.n_levels, 512),
                             s->engine.lds_share(spec.L, spec.n_levels, 1024));
        if ((int64_t)share > s->lds_budget) {
            set_error("job needs more LDS than the whole LDS budget");
            return MCRG_ERR_OUT_OF_BUDGET;
        }
        if (!s->fatal.empty()) {
            set_error("server stopped after a device error: " + s->fatal);
            return MCRG_ERR_DEVICE;
        }
        const uint64_t id = s->next_id++;
        ServerJob &j = s->jobs[id];
        j.spec = spec;
        j.t_submit = s->now();
        s->queue.push_back(id);
        s->st.submitted++;
        s->st.queued = (int64_t)s->queue.size();
        *job_id = id;
    }
    s->cv.notify_all();
    return MCRG_OK;
}

mcrg_status mcrg_server_poll(mcrg_server s, uint64_t job_id, mcrg_job_status *out)
{
    MCRG_REQUIRE(s != nullptr && out != nullptr, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        set_error("unknown job id");
        return MCRG_ERR_NOT_FOUND;
    }
    const ServerJob &j = it->second;
    out->state = j.state;
    out->n_bins = j.n_bins;
    out->submit_s = j.t_submit;
    out->start_s = j.t_start;
    out->finish_s = j.t_finish;
    return MCRG_OK;
}

mcrg_status mcrg_server_records(mcrg_server s, uint64_t job_id, int64_t *host, int64_t n)
{
    MCRG_REQUIRE(s != nullptr && host != nullptr, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        set_error("unknown job id");
        return MCRG_ERR_NOT_FOUND;
    }
    const ServerJob &j = it->second;
    if (j.state != MCRG_JOB_DONE) {
        set_error("job not finished");
        return MCRG_ERR_NOT_READY;
    }
    MCRG_REQUIRE(n == (int64_t)j.records.size(), "n must be n_bins * mcrg_record_size(n_levels)");
    std::copy(j.records.begin(), j.records.end(), host);
    return MCRG_OK;
}

mcrg_status mcrg_server_release(mcrg_server s, uint64_t job_id)
{
    MCRG_REQUIRE(s != nullptr, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        set_error("unknown job id");
        return MCRG_ERR_NOT_FOUND;
    }
    ServerJob &j = it->second;
    if (j.state == MCRG_JOB_DONE || j.state == MCRG_JOB_CANCELLED) {
        s->jobs.erase(it);
    } else {
        j.cancel = true;  // the scheduler drops it before its next launch
    }
    return MCRG_OK;
}

mcrg_status mcrg_server_completed(mcrg_server s, uint64_t *ids, int32_t max_ids, int32_t *n)
{
    MCRG_REQUIRE(s != nullptr && n != nullptr && (ids != nullptr || max_ids == 0), "null argument");
    MCRG_REQUIRE(max_ids >= 0, "max_ids must be >= 0");
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t k = 0;
    while (k < max_ids && !s->completed.empty()) {
        ids[k++] = s->completed.front();
        s->completed.pop_front();
    }
    *n = k;
    return MCRG_OK;
}

mcrg_status mcrg_server_start(mcrg_server s)
{
    MCRG_REQUIRE(s != nullptr, "null argument");
    if (s->started) return MCRG_OK;
    s->stop = false;
    s->th = std::thread(server_loop, s);
    s->started = true;
    return MCRG_OK;
}

mcrg_status mcrg_server_stop(mcrg_server s)
{
    MCRG_REQUIRE(s != nullptr, "null argument");
    if (!s->started) return MCRG_OK;
    s->stop = true;
    s->cv.notify_all();
    s->th.join();
    s->started = false;
    return MCRG_OK;
}

mcrg_status mcrg_server_step(mcrg_server s, int32_t *active)
{
    MCRG_REQUIRE(s != nullptr, "null argument");
    MCRG_REQUIRE(!s->started, "the scheduler thread is running");
    MCRG_HIP(hipSetDevice(s->ctx->device));
    return server_step(s, active);
}

mcrg_status mcrg_server_stats_get(mcrg_server s, mcrg_server_stats *out)
{
    MCRG_REQUIRE(s != nullptr && out != nullptr, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return MCRG_OK;
}

}  // extern "C"
