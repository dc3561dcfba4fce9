// Host runtime internals shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcrg_abi.h"
#include "hbm_kernels.hpp"

namespace mcrg {

void set_error(const std::string &msg);

#define MCRG_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::mcrg::set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
            return MCRG_ERR_DEVICE;                                                           \
        }                                                                                     \
    } while (0)

#define MCRG_REQUIRE(cond, msg)                                                               \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ::mcrg::set_error(msg);                                                           \
            return MCRG_ERR_INVALID_ARG;                                                      \
        }                                                                                     \
    } while (0)

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
// One bin's int64 entries cannot overflow: sweeps_per_bin (2L^2)^2 < 2^63.
inline bool bin_fits(int L, int sweeps_per_bin) { return (int64_t)sweeps_per_bin <= mcrg_max_sweeps_per_bin(L); }
#define MCRG_BIN_LIMIT_MSG "sweeps_per_bin (2 L^2)^2 must stay below 2^63 (see mcrg_max_sweeps_per_bin)"
inline int ilog2(int64_t v)
{
    int r = 0;
    while ((int64_t(1) << (r + 1)) <= v) ++r;
    return r;
}

Accept make_accept(double coupling);
// Swendsen-Wang bond threshold floor((1 - exp(-2J)) 2^32).
uint32_t bond_threshold(double coupling);

enum KernelKind { kKindSweep = 0, kKindMeasure0 = 1, kKindLevels = 2, kKindResident = 3, kKindStrip = 4, kNumKinds = 5 };

// Event-pair timing of kernel launches, per context.
struct KernelTimer {
    bool enabled = false;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int64_t launches[kNumKinds] = {0, 0, 0, 0, 0};
    double total_ms[kNumKinds] = {0, 0, 0, 0, 0};
    std::mutex mu;

    hipEvent_t get();
    // Records an event before a launch; returns the pending index or -1.
    int begin(int kind, hipStream_t s);
    void end(int idx, hipStream_t s);
    hipError_t collect();
};

}  // namespace mcrg

struct mcrg_ctx_s {
    int device = 0;
    mcrg::KernelTimer timer;
};
