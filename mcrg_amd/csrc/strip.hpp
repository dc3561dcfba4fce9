// Strip schedule for one large int8 chain: the lattice is split into n_wg
// horizontal strips of R rows, one per workgroup of a cooperative launch, and
// each strip stays bit-packed (checkerboard words) in its workgroup's LDS for
// all the sweeps of the launch.  After each half-sweep a workgroup publishes
// the first and last rows of the colour it updated and waits for its two
// neighbours' rows (point-to-point flags in device memory, no grid barrier);
// strips of at most 512 colour-1 words hold two halo rows a side instead and
// exchange once per sweep (halo).
// Measured sweeps form the level-0 operators and the blocked levels that lie
// inside the strip (R = 2^lam rows block lam times without leaving it), in the
// strip's workgroup or, with split, in a measuring workgroup per strip that
// copies each sample from it.  The bonds between a strip's last blocked row
// and the next strip's first one are counted by the upper strip later: the
// lower strip's first rows of that sample travel with a later exchange.  Each
// sample's counts and each strip's level-lam row (one row per strip) then go
// through a ring of slots to `n_tail` tail workgroups of the same launch,
// which sum the counts, measure the levels beyond lam and add the sample's
// record terms.
//
// Every word that crosses workgroups travels as one 64-bit store of (data,
// sequence number) with agent scope, and the reader polls that word until its
// sequence matches: the word carries its own validity, so no release or
// acquire fence (an L2 writeback or invalidate on this GPU) is needed between
// data and flags.
// Shared by the host runner and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hbm_kernels.hpp"

namespace mcrg {

constexpr uint32_t kStripRing = 8;         // per-sample slots between strips and tails
constexpr uint32_t kStripFlagStride = 32;  // words between polled words (128 B lines)
constexpr uint32_t kStripCntWords = 64;    // per-slot operator counts, 4 per level
// A wait gives up (and sets the error word) after 2 s of the 100 MHz clock.
constexpr uint64_t kStripWaitTicks = 200000000ull;
constexpr uint32_t kStripMaxWg = 256;
// Slots per strip of a sample's first blocked rows with halo rows: they are
// read a sweep after they are sent, while the sender may already send the
// next samples' (up to two more; up to kStripSnapSlots + 2 more when
// measuring workgroups run behind their strips).
constexpr uint32_t kStripFirstSlots = 4, kStripFirstSlotsSplit = 8;
// Sample copies a strip may publish ahead of its measuring workgroup.
constexpr uint32_t kStripSnapSlots = 4;
// LDS scratch of a strip: the consumed counts seen (kStripRing words), the
// answer of the ring-slot wait, and the word of strip_any's votes.
constexpr uint32_t kRedWords = 72, kRedWait = 68, kRedAny = 69;

struct StripShape {
    uint32_t L, nl, n_wg, R, wpr, lam, n_tail, measure;
    uint32_t threads;   // workgroup size
    // measured launches whose strips have at most one word per thread: each
    // sample is measured from a copy of the spins during the next sweep's
    // colour-0 exchange (the waits of small strips are longer than their work)
    uint32_t overlap;
    // two halo rows on each side (rows y0-2 .. y0+R+1 in LDS) and one exchange
    // per sweep: a strip updates colour 0 on its own rows, receives the
    // neighbours' two nearest colour-0 rows, then updates colour 1 on its rows
    // and the first halo row on each side itself (the same counters give the
    // neighbours' decisions).  Strips of at most 512 colour-1 words.
    uint32_t halo;
    // measured halo launches with a second workgroup per strip (split): the
    // strip publishes each sample's rows y0-1 .. y0+R (snap_words (data, seq)
    // words) and workgroup n_wg + w measures them, so a strip's sweeps no
    // longer wait for its measurements
    uint32_t split, n_meas, snap_words, first_slots;
    // device workspace, offsets in 32-bit words (even where 64-bit words
    // start), all of it cleared per launch:
    //   o_err       error word
    //   o_consumed  per slot: samples the tails have finished with it
    //   o_xhalo     (data, seq) words: [colour][strip][first, last row][wpr],
    //               with halo [sweep parity][strip][first two, last two rows][wpr]
    //   o_xfirst    (data, seq) words: [strip][first rows of levels 1..lam],
    //               with halo [sample mod first_slots][strip][...]
    //   o_ring      (data, seq) words: [slot][strip][per_wg]
    //   o_snapc     with split, per strip: samples its measuring workgroup
    //               has copied (kStripFlagStride apart)
    //   o_snap      with split, (data, seq) words: [slot][strip][snap_words]
    uint32_t o_err, o_consumed, o_xhalo, o_xfirst, o_ring, o_snapc, o_snap, ws_words;
    uint32_t cnt_words;                 // 4 counts per level 0..lam
    uint32_t first_words;               // one row of each level 1..lam
    uint32_t row_off[kMaxLevels + 1];   // level l's row within those
    uint32_t lam_words;                 // words of a level-lam row sent to the tail (0 when nl == lam)
    uint32_t per_wg;                    // cnt_words + lam_words
    // LDS, in 32-bit words
    uint32_t lds_strip, lds_tail, lds_words;
};

// Words per row of blocked level l >= 1 (row-major bits, L >> l >= 32).
__host__ __device__ inline uint32_t strip_row_words(uint32_t L, uint32_t l) { return (L >> l) >> 5; }

// n_wg strips of R = L / n_wg >= 2 rows; lam = min(log2 R, nl) levels block
// inside a strip.  `halo_ok` allows the halo-row mode where the strip fits it,
// `split_ok` its measuring workgroups.
__host__ __device__ inline StripShape strip_shape(uint32_t L, uint32_t nl, uint32_t n_wg, uint32_t n_tail,
                                                  bool measure, bool halo_ok = true, bool split_ok = false)
{
    StripShape s;
    s.L = L;
    s.nl = nl;
    s.n_wg = n_wg;
    s.R = L / n_wg;
    s.wpr = L >> 6;
    s.n_tail = measure ? n_tail : 0u;
    s.measure = measure ? 1u : 0u;
    // enough lanes for the strip's words per colour, at most 512 when tail
    // workgroups share compute units with strips
    const uint32_t words = s.R * s.wpr, hwords = (s.R + 2u) * s.wpr;
    s.halo = halo_ok && hwords <= 512u ? 1u : 0u;
    if (s.halo && measure && split_ok && words <= 256u && hwords <= 512u) {
        // measuring workgroups: 2 n_wg + n_tail workgroups resident at once,
        // three on some compute units, of at most four waves; threads j <
        // hwords - 256 update a second colour-1 word
        s.threads = hwords <= 64u ? 64u : 256u;
        s.overlap = 0u;
    } else if (s.halo) {
        s.threads = hwords <= 64u ? 64u : hwords <= 256u ? 256u : 512u;
        s.overlap = 0u;
        split_ok = false;
    } else {
        s.threads = words <= 64u ? 64u : words <= 256u ? 256u : (words <= 2048u || measure) ? 512u : 1024u;
        s.overlap = measure && words <= s.threads ? 1u : 0u;
    }
    const uint32_t rows = s.R + 2u + 2u * s.halo;  // rows held in LDS
    s.split = s.halo && measure && split_ok ? 1u : 0u;
    s.n_meas = s.split ? n_wg : 0u;
    s.snap_words = s.split ? 2u * (s.R + 2u) * s.wpr : 0u;
    s.first_slots = s.halo ? (s.split ? kStripFirstSlotsSplit : kStripFirstSlots) : 1u;
    uint32_t lg = 0;
    while ((1u << (lg + 1u)) <= s.R) ++lg;
    s.lam = measure ? (lg < nl ? lg : nl) : 0u;
    s.first_words = 0;
    for (uint32_t l = 0; l <= kMaxLevels; ++l) s.row_off[l] = 0;
    for (uint32_t l = 1; l <= s.lam; ++l) {
        s.row_off[l] = s.first_words;
        s.first_words += strip_row_words(L, l);
    }
    s.cnt_words = measure ? 4u * (s.lam + 1u) : 0u;
    s.lam_words = measure && nl > s.lam ? strip_row_words(L, s.lam) : 0u;
    s.per_wg = s.cnt_words + s.lam_words;
    s.o_err = 0;
    s.o_consumed = kStripFlagStride;
    s.o_xhalo = s.o_consumed + kStripRing * kStripFlagStride;
    s.o_xfirst = s.o_xhalo + 2u * (2u * n_wg * (2u + 2u * s.halo) * s.wpr);
    s.o_ring = s.o_xfirst + 2u * s.first_slots * n_wg * s.first_words;
    s.o_snapc = s.o_ring + (measure ? 2u * kStripRing * n_wg * s.per_wg : 0u);
    s.o_snap = s.o_snapc + (s.split ? n_wg * kStripFlagStride : 0u);
    s.ws_words = s.o_snap + 2u * kStripSnapSlots * n_wg * s.snap_words;
    // strip: both colours of rows y0-1 .. y0+R (y0-2 .. y0+R+1 with halo), the
    // in-strip levels and their tie coins, the pending sample's first and last
    // rows, the lower strip's first rows, the pending counts, reduction
    // scratch, and with overlap or halo in measured launches the copy of both
    // colours the pending sample is measured from
    uint32_t lv = 0;
    for (uint32_t l = 1; l <= s.lam; ++l) lv += (s.R >> l) * strip_row_words(L, l);
    s.lds_strip = 2u * rows * s.wpr + 2u * lv + 3u * s.first_words + kStripCntWords + kRedWords +
                  (s.overlap || (s.halo && measure) ? 2u * rows * s.wpr : 0u);
    // tail: operator values (int64) and M, the counts, reduction scratch, the
    // gathered level lam and the next one
    s.lds_tail = 0;
    if (measure) {
        const uint32_t side = L >> s.lam;
        const uint32_t a = nl > s.lam ? (side * side) / 32u : 0u;
        s.lds_tail = 2u * 3u * (kMaxLevels + 1u) + 2u + kStripCntWords + kRedWords + a + a / 4u + 1u;
    }
    s.lds_words = s.lds_strip > s.lds_tail ? s.lds_strip : s.lds_tail;
    return s;
}

struct StripArgs {
    const HbmLattice *lat;  // the chain's descriptor (device memory)
    uint32_t *ws;           // workspace of strip_shape(...).ws_words words
    StripShape sh;
    uint32_t t0;            // sweep index of the first sweep of the launch
    uint32_t q0;            // measured samples before the launch
    uint32_t n_sweeps;
    uint32_t strip_sleep;   // polls of a waiting strip thread are 64 x this many clocks apart
    uint32_t tail_sleep;    // the same for tail threads
};

// Clears the workspace, then runs n_sweeps sweeps (all measured when
// sh.measure) as one cooperative launch of sh.n_wg + sh.n_tail workgroups of
// `threads` threads, and finally stores the sweep and sample counters.
hipError_t launch_strip(const StripArgs &a, int threads, hipStream_t s);
// Workgroups of `threads` threads with `lds_bytes` of LDS resident on the
// device at once (0 on error).
int strip_max_resident(int threads, uint32_t lds_bytes);
// Workgroup size of a shape.
inline int strip_threads(const StripShape &s) { return (int)s.threads; }

}  // namespace mcrg

namespace mcrg {

struct KernelTimer;

// The allocation-free strip launches share one device array per device of
// kStaticScratchWords 32-bit words (4 MiB): enough for every shape the strip
// schedule runs (at most about 2.1 MiB, L = 8192 measured and L = 16384).
constexpr size_t kStaticScratchWords = size_t(1) << 20;
constexpr int kMaxDevices = 64;
// Looks up the array's address on `device` (the current device), once per
// device; later calls return the first answer.
hipError_t static_scratch_setup(int device);
// Its address on `device`, or null when setup failed or has not run.
uint32_t *static_scratch(int device);

// Runs a chain's sweeps on the strip schedule when its shape allows it: the
// Metropolis sampler, 64 <= L <= 8192 (L = 16384 for unmeasured sweeps, whose
// strips alone fill the LDS), and every workgroup of the launch resident at
// once.  Owns the launch workspace.  MCRG_STRIP=0 in the environment keeps
// every chain on the per-half-sweep HBM schedule (same results).
class StripRunner {
  public:
    StripRunner() = default;
    StripRunner(const StripRunner &) = delete;
    StripRunner &operator=(const StripRunner &) = delete;
    ~StripRunner();
    bool usable(const HbmLattice &h, bool measure);
    // Workspace words a launch of this shape needs (0 when it cannot run
    // here); with max_words > 0, of a launch that fits in that many.
    size_t workspace_words(const HbmLattice &h, bool measure, size_t max_words = 0);
    // With fixed_ws the launch uses that workspace of fixed_words words and
    // allocates nothing (hipErrorInvalidValue when it needs more).
    hipError_t run(const HbmLattice *d, const HbmLattice &h, uint32_t t0, uint32_t q0, uint32_t n, bool measure,
                   hipStream_t s, KernelTimer *tm, uint32_t *fixed_ws = nullptr, size_t fixed_words = 0);
    // After the stream has drained: true when a launch since the last call
    // gave up on a wait (its error word is set, then cleared here).
    hipError_t failed(hipStream_t s, bool &out);

  private:
    // The launch shape; max_words > 0 drops the measuring workgroups when
    // their workspace would not fit in that many words.
    bool plan(const HbmLattice &h, bool measure, StripShape &sh, int &threads, size_t max_words);
    uint32_t *ws_ = nullptr;
    size_t ws_words_ = 0;
    uint32_t *last_ws_ = nullptr;  // workspace of the last launch (its error word)
    bool launched_ = false;
    int cus_ = -1;
};

}  // namespace mcrg
