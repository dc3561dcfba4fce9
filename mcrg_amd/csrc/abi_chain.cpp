// C-ABI: context, chain (HBM schedule) and timing entry points.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "hbm_group.hpp"
#include "runtime.hpp"
#include "strip.hpp"

namespace mcrg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

Accept make_accept(double coupling)
{
    uint64_t thr[2];
    mcrg_thresholds(coupling, thr);
    Accept a;
    a.always0 = thr[0] >= (uint64_t(1) << 32) ? 1u : 0u;
    a.always1 = thr[1] >= (uint64_t(1) << 32) ? 1u : 0u;
    a.thr0 = a.always0 ? 0xFFFFFFFFu : (uint32_t)thr[0];
    a.thr1 = a.always1 ? 0xFFFFFFFFu : (uint32_t)thr[1];
    return a;
}

uint32_t bond_threshold(double coupling)
{
    double v = std::floor(-std::expm1(-2.0 * coupling) * 4294967296.0);
    if (v < 0.0) v = 0.0;
    if (v > 4294967295.0) v = 4294967295.0;
    return (uint32_t)v;
}

hipEvent_t KernelTimer::get()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int KernelTimer::begin(int kind, hipStream_t s)
{
    if (!enabled) return -1;
    std::lock_guard<std::mutex> lk(mu);
    Pending p{kind, get(), get()};
    if (!p.a || !p.b) return -1;
    if (hipEventRecord(p.a, s) != hipSuccess) return -1;
    pending.push_back(p);
    return (int)pending.size() - 1;
}

void KernelTimer::end(int idx, hipStream_t s)
{
    if (idx < 0) return;
    std::lock_guard<std::mutex> lk(mu);
    (void)hipEventRecord(pending[(size_t)idx].b, s);
}

hipError_t KernelTimer::collect()
{
    std::lock_guard<std::mutex> lk(mu);
    for (auto &p : pending) {
        hipError_t e = hipEventSynchronize(p.b);
        if (e != hipSuccess) return e;
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, p.a, p.b);
        if (e != hipSuccess) return e;
        launches[p.kind] += 1;
        total_ms[p.kind] += ms;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    return hipSuccess;
}

}  // namespace mcrg

using namespace mcrg;

// ------------------------------------------------------------------ chain object
struct mcrg_chain_s {
    mcrg_ctx ctx = nullptr;
    mcrg_chain_desc desc{};
    hipStream_t stream = nullptr;
    HbmLattice host{};            // descriptor (device pointers), host copy
    HbmLattice *d_desc = nullptr;  // device copy
    void *d_block = nullptr;       // one allocation for spins, level-1, sums, state, records
    uint32_t t = 0;                // host mirror of state[0]
    uint32_t n_measured = 0;       // host mirror of state[1]
    SweepRunner runner;            // replayed graphs of long sweep runs
    StripRunner strip;             // cooperative strip launches (large Metropolis chains)
};

// Fails when a strip launch since the last check gave up on a wait.
static mcrg_status check_strip(mcrg_chain ch)
{
    bool bad = false;
    MCRG_HIP(ch->strip.failed(ch->stream, bad));
    if (bad) {
        set_error("strip launch: a workgroup's wait for its neighbours exceeded the limit; the chain state is lost");
        return MCRG_ERR_DEVICE;
    }
    return MCRG_OK;
}

extern "C" {

int32_t mcrg_abi_version(void) { return MCRG_ABI_VERSION; }

const char *mcrg_last_error(void) { return g_last_error.c_str(); }

mcrg_status mcrg_thresholds(double coupling, uint64_t thr[2])
{
    MCRG_REQUIRE(thr != nullptr, "thr is null");
    MCRG_REQUIRE(std::isfinite(coupling) && coupling >= 0.0,
                 "coupling J = |K| must be finite and >= 0 (ferromagnetic; reference K = -J)");
    for (int q = 0; q < 2; ++q) {
        const double p = std::exp(-4.0 * coupling * (double)(q + 1));
        double v = std::floor(p * 4294967296.0);
        if (v < 0.0) v = 0.0;
        if (v > 4294967296.0) v = 4294967296.0;
        thr[q] = (uint64_t)v;
    }
    return MCRG_OK;
}

mcrg_status mcrg_bond_threshold(double coupling, uint32_t *thr)
{
    MCRG_REQUIRE(thr != nullptr, "thr is null");
    MCRG_REQUIRE(std::isfinite(coupling) && coupling >= 0.0,
                 "coupling J = |K| must be finite and >= 0 (ferromagnetic; reference K = -J)");
    *thr = bond_threshold(coupling);
    return MCRG_OK;
}

int32_t mcrg_record_size(int32_t n_levels)
{
    if (n_levels < 0 || n_levels > (int32_t)kMaxLevels) return -1;
    return (int32_t)record_size((uint32_t)n_levels);
}

mcrg_status mcrg_ctx_create(int32_t device, mcrg_ctx *out)
{
    MCRG_REQUIRE(out != nullptr, "out is null");
    int n = 0;
    MCRG_HIP(hipGetDeviceCount(&n));
    MCRG_REQUIRE(device >= 0 && device < n, "device index out of range");
    MCRG_HIP(hipSetDevice(device));
    mcrg_ctx c = new (std::nothrow) mcrg_ctx_s;
    if (!c) {
        set_error("out of host memory");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    c->device = device;
    // the allocation-free sweep's workspace on this device (see strip.hpp); a
    // failure is reported by that entry point, not here
    (void)static_scratch_setup(device);
    *out = c;
    return MCRG_OK;
}

mcrg_status mcrg_ctx_destroy(mcrg_ctx ctx)
{
    if (!ctx) return MCRG_OK;
    (void)hipSetDevice(ctx->device);
    (void)ctx->timer.collect();
    for (hipEvent_t e : ctx->timer.pool) (void)hipEventDestroy(e);
    delete ctx;
    return MCRG_OK;
}

mcrg_status mcrg_ctx_device_info(mcrg_ctx ctx, mcrg_device_info *out)
{
    MCRG_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
    hipDeviceProp_t p;
    MCRG_HIP(hipGetDeviceProperties(&p, ctx->device));
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.gcnArchName, sizeof(out->name) - 1);
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->lds_per_block = (int32_t)p.sharedMemPerBlock;
    out->warp_size = p.warpSize;
    out->global_mem_bytes = (int64_t)p.totalGlobalMem;
    return MCRG_OK;
}

mcrg_status mcrg_ctx_set_timing(mcrg_ctx ctx, int32_t enable)
{
    MCRG_REQUIRE(ctx != nullptr, "ctx is null");
    ctx->timer.enabled = enable != 0;
    return MCRG_OK;
}

mcrg_status mcrg_ctx_kernel_time(mcrg_ctx ctx, int32_t kind, int64_t *launches, double *total_ms)
{
    MCRG_REQUIRE(ctx != nullptr, "ctx is null");
    MCRG_REQUIRE(kind >= 0 && kind < kNumKinds, "kernel kind out of range");
    MCRG_HIP(hipSetDevice(ctx->device));
    MCRG_HIP(ctx->timer.collect());
    if (launches) *launches = ctx->timer.launches[kind];
    if (total_ms) *total_ms = ctx->timer.total_ms[kind];
    return MCRG_OK;
}

static size_t align256(size_t v) { return (v + 255u) & ~size_t(255u); }

mcrg_status mcrg_chain_create(mcrg_ctx ctx, const mcrg_chain_desc *desc, mcrg_chain *out)
{
    MCRG_REQUIRE(ctx != nullptr && desc != nullptr && out != nullptr, "null argument");
    const mcrg_chain_desc d = *desc;
    MCRG_REQUIRE(is_pow2(d.L) && d.L >= 2 && d.L <= 16384, "L must be a power of two in [2, 16384]");
    MCRG_REQUIRE(std::isfinite(d.coupling) && d.coupling >= 0.0, "coupling J = |K| must be >= 0");
    MCRG_REQUIRE(d.n_levels >= 0 && d.n_levels <= ilog2(d.L) - 1 && d.n_levels <= (int)kMaxLevels,
                 "n_levels must be in [0, log2(L) - 1]");
    MCRG_REQUIRE(d.sweeps_per_bin >= 1 && d.max_bins >= 1, "sweeps_per_bin and max_bins must be >= 1");
    MCRG_REQUIRE(bin_fits(d.L, d.sweeps_per_bin), MCRG_BIN_LIMIT_MSG);
    MCRG_REQUIRE(d.layout == MCRG_LAYOUT_INT8, "only MCRG_LAYOUT_INT8 chains run on this schedule");
    MCRG_REQUIRE(d.sampler == MCRG_SAMPLER_METROPOLIS || d.sampler == MCRG_SAMPLER_SWENDSEN_WANG,
                 "sampler must be MCRG_SAMPLER_METROPOLIS or MCRG_SAMPLER_SWENDSEN_WANG");
    MCRG_HIP(hipSetDevice(ctx->device));

    mcrg_chain ch = new (std::nothrow) mcrg_chain_s;
    if (!ch) {
        set_error("out of host memory");
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    ch->ctx = ctx;
    ch->desc = d;
    const HbmBlock blk = hbm_block(d.L, d.n_levels, d.max_bins, d.sampler);
    const size_t o_desc = blk.bytes;
    const size_t total = align256(o_desc + sizeof(HbmLattice));
    hipError_t e = hipMalloc(&ch->d_block, total);
    if (e != hipSuccess) {
        delete ch;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return MCRG_ERR_OUT_OF_BUDGET;
    }
    char *base = static_cast<char *>(ch->d_block);
    ch->host = hbm_descriptor(base, blk, d.L, d.n_levels, d.sweeps_per_bin, d.seed, d.coupling, d.sampler);
    ch->d_desc = reinterpret_cast<HbmLattice *>(base + o_desc);

    auto fail = [&](hipError_t err, const char *what) {
        set_error(std::string(what) + ": " + hipGetErrorString(err));
        if (ch->stream) (void)hipStreamDestroy(ch->stream);
        (void)hipFree(ch->d_block);
        delete ch;
        return MCRG_ERR_DEVICE;
    };
    if ((e = hipStreamCreateWithFlags(&ch->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreateWithFlags");
    if ((e = hipMemsetAsync(ch->d_block, 0, o_desc, ch->stream)) != hipSuccess) return fail(e, "hipMemsetAsync");
    if ((e = hipMemcpyAsync(ch->d_desc, &ch->host, sizeof(HbmLattice), hipMemcpyHostToDevice,
                            ch->stream)) != hipSuccess)
        return fail(e, "hipMemcpyAsync");
    if ((e = launch_int8_init(ch->d_desc, 1, d.L, ch->stream)) != hipSuccess) return fail(e, "k_int8_init");
    if ((e = hipStreamSynchronize(ch->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    *out = ch;
    return MCRG_OK;
}

mcrg_status mcrg_chain_destroy(mcrg_chain ch)
{
    if (!ch) return MCRG_OK;
    (void)hipSetDevice(ch->ctx->device);
    (void)hipStreamSynchronize(ch->stream);
    (void)hipStreamDestroy(ch->stream);
    (void)hipFree(ch->d_block);
    delete ch;
    return MCRG_OK;
}

mcrg_status mcrg_chain_set_coupling(mcrg_chain ch, double coupling)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    MCRG_REQUIRE(std::isfinite(coupling) && coupling >= 0.0, "coupling J = |K| must be >= 0");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    ch->desc.coupling = coupling;
    ch->host.acc = make_accept(coupling);
    ch->host.bond_thr = bond_threshold(coupling);
    MCRG_HIP(hipMemcpyAsync(ch->d_desc, &ch->host, sizeof(HbmLattice), hipMemcpyHostToDevice, ch->stream));
    // the host descriptor may change again before the copy runs
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    return MCRG_OK;
}

mcrg_status mcrg_chain_sweep(mcrg_chain ch, int32_t n_sweeps, int32_t measure)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    MCRG_REQUIRE(n_sweeps >= 0, "n_sweeps must be >= 0");
    if (n_sweeps == 0) return MCRG_OK;
    const mcrg_chain_desc &d = ch->desc;
    if (measure) {
        const uint64_t last = (uint64_t)ch->n_measured + (uint64_t)n_sweeps - 1;
        if (last / (uint64_t)d.sweeps_per_bin >= (uint64_t)d.max_bins) {
            set_error("record bins exhausted: raise max_bins or sweeps_per_bin");
            return MCRG_ERR_CAPACITY;
        }
    }
    MCRG_REQUIRE((uint64_t)ch->t + (uint64_t)n_sweeps <= 0xFFFFFFFFull, "sweep counter overflow");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    KernelTimer *tm = ch->ctx->timer.enabled ? &ch->ctx->timer : nullptr;
    if (ch->strip.usable(ch->host, measure != 0))
        MCRG_HIP(ch->strip.run(ch->d_desc, ch->host, ch->t, ch->n_measured, (uint32_t)n_sweeps, measure != 0,
                               ch->stream, tm));
    else
        MCRG_HIP(ch->runner.run(ch->d_desc, 1, d.L, d.n_levels, d.sampler, (uint32_t)n_sweeps, measure != 0,
                                ch->stream, tm));
    ch->t += (uint32_t)n_sweeps;
    if (measure) ch->n_measured += (uint32_t)n_sweeps;
    return MCRG_OK;
}

mcrg_status mcrg_chain_sweep_static(mcrg_chain ch, int32_t n_sweeps, int32_t measure)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    MCRG_REQUIRE(n_sweeps >= 0, "n_sweeps must be >= 0");
    if (n_sweeps == 0) return MCRG_OK;
    const mcrg_chain_desc &d = ch->desc;
    if (measure) {
        const uint64_t last = (uint64_t)ch->n_measured + (uint64_t)n_sweeps - 1;
        if (last / (uint64_t)d.sweeps_per_bin >= (uint64_t)d.max_bins) {
            set_error("record bins exhausted: raise max_bins or sweeps_per_bin");
            return MCRG_ERR_CAPACITY;
        }
    }
    MCRG_REQUIRE((uint64_t)ch->t + (uint64_t)n_sweeps <= 0xFFFFFFFFull, "sweep counter overflow");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    uint32_t *ws = static_scratch(ch->ctx->device);
    if (!ws) {
        set_error("the static workspace of this device could not be located (hipGetSymbolAddress failed at context "
                  "creation)");
        return MCRG_ERR_DEVICE;
    }
    const size_t need = ch->strip.workspace_words(ch->host, measure != 0, kStaticScratchWords);
    MCRG_REQUIRE(need > 0, "mcrg_chain_sweep_static runs chains on the strip schedule only: the Metropolis sampler, "
                           "64 <= L <= 8192 (L = 16384 for unmeasured sweeps)");
    MCRG_REQUIRE(need <= kStaticScratchWords, "this chain needs more workspace than the static array holds");
    MCRG_HIP(ch->strip.run(ch->d_desc, ch->host, ch->t, ch->n_measured, (uint32_t)n_sweeps, measure != 0,
                           ch->stream, nullptr, ws, kStaticScratchWords));
    ch->t += (uint32_t)n_sweeps;
    if (measure) ch->n_measured += (uint32_t)n_sweeps;
    return MCRG_OK;
}

mcrg_status mcrg_chain_sync(mcrg_chain ch)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    return check_strip(ch);
}

mcrg_status mcrg_chain_get_spins(mcrg_chain ch, int8_t *host, int64_t n)
{
    MCRG_REQUIRE(ch != nullptr && host != nullptr, "null argument");
    MCRG_REQUIRE(n == (int64_t)ch->desc.L * ch->desc.L, "n must be L*L");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    MCRG_HIP(hipMemcpyAsync(host, ch->host.spins, (size_t)n, hipMemcpyDeviceToHost, ch->stream));
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    return check_strip(ch);
}

mcrg_status mcrg_chain_set_spins(mcrg_chain ch, const int8_t *host, int64_t n)
{
    MCRG_REQUIRE(ch != nullptr && host != nullptr, "null argument");
    MCRG_REQUIRE(n == (int64_t)ch->desc.L * ch->desc.L, "n must be L*L");
    for (int64_t i = 0; i < n; ++i)
        MCRG_REQUIRE(host[i] == 1 || host[i] == -1, "spins must be +1 or -1");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    MCRG_HIP(hipMemcpyAsync(ch->host.spins, host, (size_t)n, hipMemcpyHostToDevice, ch->stream));
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    return MCRG_OK;
}

mcrg_status mcrg_chain_get_records(mcrg_chain ch, int64_t *host, int64_t n)
{
    MCRG_REQUIRE(ch != nullptr && host != nullptr, "null argument");
    const int64_t want = (int64_t)ch->desc.max_bins * record_size((uint32_t)ch->desc.n_levels);
    MCRG_REQUIRE(n == want, "n must be max_bins * mcrg_record_size(n_levels)");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    MCRG_HIP(hipMemcpyAsync(host, ch->host.records, (size_t)n * 8, hipMemcpyDeviceToHost, ch->stream));
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    return check_strip(ch);
}

mcrg_status mcrg_chain_get_state(mcrg_chain ch, uint32_t *sweep, uint32_t *n_measured)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    if (sweep) *sweep = ch->t;
    if (n_measured) *n_measured = ch->n_measured;
    return MCRG_OK;
}

mcrg_status mcrg_chain_set_state(mcrg_chain ch, uint32_t sweep, uint32_t n_measured,
                                 const int64_t *records, int64_t n)
{
    MCRG_REQUIRE(ch != nullptr, "chain is null");
    const int64_t want = (int64_t)ch->desc.max_bins * record_size((uint32_t)ch->desc.n_levels);
    MCRG_REQUIRE(records == nullptr || n == want, "n must be max_bins * mcrg_record_size(n_levels)");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    const uint32_t st[2] = {sweep, n_measured};
    MCRG_HIP(hipMemcpyAsync(ch->host.state, st, sizeof(st), hipMemcpyHostToDevice, ch->stream));
    if (records)
        MCRG_HIP(hipMemcpyAsync(ch->host.records, records, (size_t)n * 8, hipMemcpyHostToDevice, ch->stream));
    MCRG_HIP(hipStreamSynchronize(ch->stream));
    ch->t = sweep;
    ch->n_measured = n_measured;
    return MCRG_OK;
}

mcrg_status mcrg_chain_operators(mcrg_chain ch, int64_t out[4])
{
    MCRG_REQUIRE(ch != nullptr && out != nullptr, "null argument");
    MCRG_HIP(hipSetDevice(ch->ctx->device));
    hipStream_t s = ch->stream;
    MCRG_HIP(hipMemsetAsync(ch->host.sums, 0, 4 * sizeof(unsigned long long), s));
    MCRG_HIP(launch_int8_measure0(ch->d_desc, 1, ch->desc.L, 0u, s));
    unsigned long long cnt[4];
    MCRG_HIP(hipMemcpyAsync(cnt, ch->host.sums, sizeof(cnt), hipMemcpyDeviceToHost, s));
    MCRG_HIP(hipMemsetAsync(ch->host.sums, 0, 4 * sizeof(unsigned long long), s));
    MCRG_HIP(hipStreamSynchronize(s));
    const int64_t n0 = (int64_t)ch->desc.L * ch->desc.L;
    out[0] = 2 * n0 - 2 * (int64_t)cnt[0];
    out[1] = 2 * n0 - 2 * (int64_t)cnt[1];
    out[2] = n0 - 2 * (int64_t)cnt[2];
    out[3] = n0 - 2 * (int64_t)cnt[3];
    return MCRG_OK;
}

}  // extern "C"
