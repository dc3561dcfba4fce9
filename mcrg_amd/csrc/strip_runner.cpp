// Host side of the strip schedule (strip.hpp).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "runtime.hpp"
#include "strip.hpp"

namespace mcrg {

// Workspace of the allocation-free strip launches (mcrg_chain_sweep_static):
// each device the library runs on holds this array until the process ends.
// In host code its name is only a stand-in for the device copy: the launches
// use the per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) uint32_t scratch[kStaticScratchWords];

namespace {
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
}  // namespace

hipError_t static_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        // the calling thread has made `device` current
        void *p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? static_cast<uint32_t *>(p) : nullptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_err[device];
}

uint32_t *static_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

StripRunner::~StripRunner()
{
    if (ws_) (void)hipFree(ws_);
}

static bool strip_enabled()
{
    const char *v = std::getenv("MCRG_STRIP");
    return !(v && std::strcmp(v, "0") == 0);
}

bool StripRunner::plan(const HbmLattice &h, bool measure, StripShape &sh, int &threads, size_t max_words)
{
    if (h.sampler != MCRG_SAMPLER_METROPOLIS || !strip_enabled()) return false;
    const uint32_t L = (uint32_t)h.L;
    if (L < 64u || L > (measure ? 8192u : 16384u)) return false;
    if (cus_ < 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus_ = 0;
    }
    uint32_t n_wg = 1;
    while (2u * n_wg <= (uint32_t)cus_ && 2u * n_wg <= kStripMaxWg) n_wg *= 2u;
    if (const char *v = std::getenv("MCRG_STRIP_WG")) {  // fewer strips (a power of two) when asked
        const uint32_t want = (uint32_t)std::max(2, std::atoi(v));
        while (n_wg > want) n_wg /= 2u;
    }
    if (n_wg > L / 2u) n_wg = L / 2u;
    if (n_wg < 2u) return false;
    // tail workgroups: 4 (fewer when they do not fit beside the strips);
    // MCRG_STRIP_TAILS = 1, 2 or 4 asks for that many
    uint32_t want = 4u;
    if (const char *v = std::getenv("MCRG_STRIP_TAILS")) {
        const int n = std::atoi(v);
        if (n == 1 || n == 2 || n == 4) want = (uint32_t)n;
    }
    // MCRG_STRIP_HALO=0 keeps two exchanges per sweep where halo rows would
    // fit; MCRG_STRIP_SPLIT=0 keeps a halo strip's measurements in its own
    // workgroup
    const char *hv = std::getenv("MCRG_STRIP_HALO");
    const bool halo_ok = !(hv && std::atoi(hv) == 0);
    const char *sv = std::getenv("MCRG_STRIP_SPLIT");
    const bool split_ok = !(sv && std::atoi(sv) == 0);
    for (uint32_t n_tail : {4u, 2u, 1u}) {
        if (n_tail > want) continue;
        // measuring workgroups when they fit (resident and within max_words)
        for (int sp = split_ok ? 1 : 0; sp >= 0; --sp) {
            sh = strip_shape(L, (uint32_t)h.n_levels, n_wg, n_tail, measure, halo_ok, sp != 0);
            if (sp && !sh.split) continue;
            threads = strip_threads(sh);
            // every thread gets at most one neighbour word of each kind per exchange
            if ((sh.halo ? 4u : 2u) * sh.wpr > (uint32_t)threads || sh.first_words > (uint32_t)threads) return false;
            const uint32_t lds = sh.lds_words * 4u;
            if (lds > 160u * 1024u) return false;
            if (max_words && sh.ws_words > max_words) continue;
            if ((uint32_t)strip_max_resident(threads, lds) >= sh.n_wg + sh.n_meas + sh.n_tail) return true;
        }
        if (!measure) return false;
    }
    return false;
}

bool StripRunner::usable(const HbmLattice &h, bool measure)
{
    StripShape sh;
    int threads = 0;
    return plan(h, measure, sh, threads, 0);
}

hipError_t StripRunner::run(const HbmLattice *d, const HbmLattice &h, uint32_t t0, uint32_t q0, uint32_t n,
                            bool measure, hipStream_t s, KernelTimer *tm, uint32_t *fixed_ws, size_t fixed_words)
{
    StripArgs a;
    int threads = 0;
    if (!plan(h, measure, a.sh, threads, fixed_ws ? fixed_words : 0)) return hipErrorInvalidValue;
    hipError_t e;
    if (fixed_ws) {
        if (a.sh.ws_words > fixed_words) return hipErrorInvalidValue;
        a.ws = fixed_ws;
    } else if (a.sh.ws_words > ws_words_) {
        // the stream may still run a launch that uses the old workspace
        if (ws_) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            (void)hipFree(ws_);
            ws_ = nullptr;
        }
        void *p = nullptr;
        if ((e = hipMalloc(&p, (size_t)a.sh.ws_words * 4u)) != hipSuccess) return e;
        ws_ = static_cast<uint32_t *>(p);
        ws_words_ = a.sh.ws_words;
    }
    if (!fixed_ws) a.ws = ws_;
    a.lat = d;
    a.t0 = t0;
    a.q0 = q0;
    a.n_sweeps = n;
    a.strip_sleep = 1;
    a.tail_sleep = 8;
    if (const char *v = std::getenv("MCRG_STRIP_SLEEP")) a.strip_sleep = (uint32_t)std::max(1, std::atoi(v));
    if (const char *v = std::getenv("MCRG_TAIL_SLEEP")) a.tail_sleep = (uint32_t)std::max(1, std::atoi(v));
    const int ev = tm ? tm->begin(kKindStrip, s) : -1;
    if ((e = launch_strip(a, threads, s)) != hipSuccess) return e;
    if (tm) tm->end(ev, s);
    launched_ = true;
    last_ws_ = a.ws;
    return hipSuccess;
}

size_t StripRunner::workspace_words(const HbmLattice &h, bool measure, size_t max_words)
{
    StripShape sh;
    int threads = 0;
    return plan(h, measure, sh, threads, max_words) ? (size_t)sh.ws_words : 0u;
}

hipError_t StripRunner::failed(hipStream_t s, bool &out)
{
    out = false;
    if (!launched_) return hipSuccess;
    uint32_t err = 0;
    hipError_t e;
    if ((e = hipMemcpyAsync(&err, last_ws_, sizeof(err), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    launched_ = false;
    out = err != 0u;
    return hipSuccess;  // the next launch clears the word
}

}  // namespace mcrg
