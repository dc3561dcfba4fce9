/*
 * mcrg_abi.h -- C ABI of the MI355X MCRG hot path (libmcrg_amd.so).
 *
 * Plain C types only: opaque handles, pointers and sizes, int status codes.
 * Nothing here throws; on failure a function returns a negative mcrg_status and
 * mcrg_last_error() describes it (thread-local).
 *
 * The reference (kim-jane/MCRG) has no FFI: its boundary is three C++ classes.
 * Each entry point below names the reference member it stands in for; the
 * C++ classes that keep the reference's names sit on top of this ABI
 * (mcrg_amd/host/).
 *
 * Model and stream conventions: coupling J = |K| > 0 is ferromagnetic (the
 * reference writes K = -J, src/ising.cpp:9); spins are +1/-1 int8 in row-major
 * order, site (x, y) at y*L + x (the reference's column-major (i, j) at j*N + i);
 * one "sweep" is a red/black checkerboard Metropolis pass driven by
 * Philox4x32-10 keyed by the seed.  Operator sums are single-counted (the
 * reference's calc_interactions returns twice NN and NNN).
 */
#ifndef MCRG_ABI_H
#define MCRG_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCRG_ABI_VERSION 3

typedef int32_t mcrg_status;
enum {
    MCRG_OK = 0,
    MCRG_ERR_INVALID_ARG = -1,   /* bad size, coupling, level count, handle ... */
    MCRG_ERR_OUT_OF_BUDGET = -2, /* device memory or admission budget exhausted */
    MCRG_ERR_DEVICE = -3,        /* a HIP call failed */
    MCRG_ERR_NOT_READY = -4,     /* result not available yet */
    MCRG_ERR_NOT_FOUND = -5,     /* unknown job id */
    MCRG_ERR_CAPACITY = -6       /* record bins exhausted */
};

enum { MCRG_LAYOUT_INT8 = 0, MCRG_LAYOUT_PACKED = 1 };

/* Samplers: one "sweep" is a red/black checkerboard Metropolis pass, or one
 * Swendsen-Wang pass (every aligned pair bonded with probability 1 - exp(-2J),
 * every cluster flipped with probability 1/2), the all-clusters form of the
 * reference's Wolff update (src/ising.cpp:87-149).  Swendsen-Wang runs on the
 * HBM schedule (int8). */
enum { MCRG_SAMPLER_METROPOLIS = 0, MCRG_SAMPLER_SWENDSEN_WANG = 1 };

typedef struct mcrg_ctx_s *mcrg_ctx;
typedef struct mcrg_chain_s *mcrg_chain;

/* Library version and last error of the calling thread. */
int32_t mcrg_abi_version(void);
const char *mcrg_last_error(void);

/* Metropolis thresholds floor(exp(-4J q) 2^32), q = 1, 2 (a value of 2^32 means
 * "always flip").  Replaces P_ = 1 - exp(2K) of IsingModel(double K),
 * src/ising.cpp:9. */
mcrg_status mcrg_thresholds(double coupling, uint64_t thr[2]);

/* Swendsen-Wang bond threshold floor((1 - exp(-2J)) 2^32): an aligned pair is
 * bonded when its 32-bit uniform is below it.  The reference's Wolff step adds
 * an aligned neighbour with P_ = 1 - exp(2K) (src/ising.cpp:9,134). */
mcrg_status mcrg_bond_threshold(double coupling, uint32_t *thr);

/* int64 entries in one per-bin record for n_levels blockings:
 * 22 n_levels + 16.  Layout (l = 0..n_levels, a, b in {NN, NNN, P}):
 *   [l*3 + a]                                 sum S_l[a]
 *   [3(nl+1) + l*9 + b*3 + a]                 sum S_l[a] S_l[b]
 *   [12(nl+1) + (l-1)*9 + b*3 + a], l >= 1    sum S_l[a] S_{l-1}[b]
 *   [12(nl+1) + 9nl + l - 1], l >= 1          sum NN_l NN_0
 *   [12(nl+1) + 10nl + 0..3]                  sum M, sum |M|, sum M^2, samples
 * These are the accumulators S_avg_loc, Sb_Sb_avg_loc, Sb_S_avg_loc of
 * calc_critical_exponent (src/mcrg.cpp:59-66, flattened as
 * src/definitions.cpp:9-19) and SLb_SL_avg_loc of approx_critical_point
 * (src/mcrg.cpp:262-263), kept exactly in int64 per bin. */
int32_t mcrg_record_size(int32_t n_levels);

/* Largest sweeps_per_bin for which one bin's int64 entries cannot overflow on a
 * side-L lattice: floor((2^63 - 1) / (2 L^2)^2), the bound of a level-0 self
 * product (2L^2 single-counted bonds, all aligned) summed over the bin's
 * samples (0 for an L out of range).  Chains, batch jobs and server jobs with
 * more sweeps per bin are rejected with MCRG_ERR_INVALID_ARG.  Sums over bins
 * are formed in 128-bit integers by the solve functions below.  No reference
 * counterpart: the reference keeps its sums in FP64 (src/mcrg.cpp:59-66). */
int64_t mcrg_max_sweeps_per_bin(int32_t L);

/* A context owns one device.  Replaces the MPI_Init / MPI_COMM_WORLD setup of
 * src/main.cpp:6 (one process per device; no MPI). */
mcrg_status mcrg_ctx_create(int32_t device, mcrg_ctx *out);
mcrg_status mcrg_ctx_destroy(mcrg_ctx ctx);

/* Properties of the context's device (no reference counterpart: the reference
 * runs on host cores only). */
typedef struct {
    char name[64];            /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t compute_units;
    int32_t clock_khz;
    int32_t lds_per_block;    /* bytes a workgroup may allocate            */
    int32_t warp_size;
    int64_t global_mem_bytes;
} mcrg_device_info;
mcrg_status mcrg_ctx_device_info(mcrg_ctx ctx, mcrg_device_info *out);

/* ------------------------------------------------------------------ chains */
typedef struct {
    int32_t L;              /* side, power of two, 2 <= L <= 16384          */
    double coupling;        /* J = |K| >= 0                                  */
    uint64_t seed;          /* Philox key of this chain                      */
    int32_t n_levels;       /* blockings per sample, 0 <= n <= log2(L) - 1   */
    int32_t layout;         /* MCRG_LAYOUT_INT8 or MCRG_LAYOUT_PACKED        */
    int32_t sweeps_per_bin; /* measured sweeps per record bin (>= 1)         */
    int32_t max_bins;       /* record bins allocated (>= 1)                  */
    int32_t sampler;        /* MCRG_SAMPLER_* (0 = Metropolis)               */
} mcrg_chain_desc;

/* Allocates the chain on the device and draws the Philox +-1 start.
 * Replaces Lattice(int N) + initialize_random_spins (src/lattice.cpp:3-15,33-41)
 * together with IsingModel(double K) (src/ising.cpp:3-11). */
mcrg_status mcrg_chain_create(mcrg_ctx ctx, const mcrg_chain_desc *desc, mcrg_chain *out);
mcrg_status mcrg_chain_destroy(mcrg_chain ch);

/* Changes the coupling for subsequent sweeps (IsingModel(double K) applied to an
 * existing Lattice, src/mcrg.cpp:46-50). */
mcrg_status mcrg_chain_set_coupling(mcrg_chain ch, double coupling);

/* Enqueues n_sweeps sweeps (asynchronous).  With measure != 0 every sweep is
 * followed by the level 0..n_levels measurement, accumulated into the record
 * of bin (sample index / sweeps_per_bin).  One sweep stands in for one
 * IsingModel::sample_new_configuration (src/ising.cpp:87-93); a run of
 * unmeasured sweeps is IsingModel::equilibrate(.., write = false)
 * (src/ising.cpp:14-84); measured sweeps are the sampling loop of
 * calc_critical_exponent (src/mcrg.cpp:72-98). */
mcrg_status mcrg_chain_sweep(mcrg_chain ch, int32_t n_sweeps, int32_t measure);

/* Allocation-free variant of mcrg_chain_sweep.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It runs the sweeps on the strip schedule (the Metropolis sampler,
 * 64 <= L <= 8192, and L = 16384 for unmeasured sweeps) with the launch's
 * workspace in a static device array of 4 MiB per device, located once per
 * device when a context is created, and calls no allocating routine; other
 * chains are rejected with MCRG_ERR_INVALID_ARG.  Spins, counters and records
 * are bit-identical to mcrg_chain_sweep's.  Every call on a device shares that
 * device's array, which keeps the last call's contents (each launch clears
 * what it uses before it starts): calls that overlap on one device race, so
 * enqueue them on one chain's stream or synchronise between them. */
mcrg_status mcrg_chain_sweep_static(mcrg_chain ch, int32_t n_sweeps, int32_t measure);

/* Waits for the chain's queued work. */
mcrg_status mcrg_chain_sync(mcrg_chain ch);

/* Spins as +-1 int8, row-major, n = L*L (synchronous).  Backs the public
 * Lattice::spins_ (src/lattice.hpp:16) and checkpoint/restore. */
mcrg_status mcrg_chain_get_spins(mcrg_chain ch, int8_t *host, int64_t n);
mcrg_status mcrg_chain_set_spins(mcrg_chain ch, const int8_t *host, int64_t n);

/* All max_bins records, n = max_bins * mcrg_record_size(n_levels) (synchronous). */
mcrg_status mcrg_chain_get_records(mcrg_chain ch, int64_t *host, int64_t n);

/* Sweep counter t (next sweep index) and number of measured samples. */
mcrg_status mcrg_chain_get_state(mcrg_chain ch, uint32_t *sweep, uint32_t *n_measured);
/* Restores counters and records (resume from a checkpoint); n as get_records. */
mcrg_status mcrg_chain_set_state(mcrg_chain ch, uint32_t sweep, uint32_t n_measured,
                                 const int64_t *records, int64_t n);

/* Single-count NN, NNN, plaquette and magnetisation of the current spins
 * (synchronous).  Lattice::calc_interactions (src/lattice.cpp:102-120) returns
 * 2*out[0], 2*out[1]; IsingModel::calc_magnetization (src/ising.cpp:176-179)
 * is out[3] / L^2. */
mcrg_status mcrg_chain_operators(mcrg_chain ch, int64_t out[4]);

/* ------------------------------------------------------------------ batches */
/* Many independent chains (uneven L and coupling) advanced together.  Jobs of
 * 8 <= L <= 512 with the Metropolis sampler run on the LDS-resident schedule:
 * bit-packed lattices held in LDS by a persistent grid, one launch per chunk of
 * sweeps.  Jobs of L = 2, 4 and L >= 1024, and Swendsen-Wang jobs, run on the
 * HBM schedule (int8 in device memory): jobs of the same L, n_levels, n_therm,
 * n_meas, sweeps_per_bin and sampler share every launch, on
 * a second stream that overlaps the resident launches, and long runs replay
 * captured HIP graphs of 32 sweeps.  A job is the reference's whole
 * calc_critical_exponent chain (src/mcrg.cpp:22-98) for one (L, K, seed):
 * n_therm unmeasured sweeps (equilibrate, src/ising.cpp:14-84) then n_meas
 * measured sweeps.  Every job is bit-identical to the same chain run alone
 * through mcrg_chain_* (same seed, same Philox counters).  The reference
 * spreads chains over MPI ranks (split_samples, src/definitions.cpp:79-87);
 * here they share one device. */
typedef struct mcrg_batch_s *mcrg_batch;

typedef struct {
    int32_t L;              /* power of two, 2 <= L <= 16384 (server: 8..512)  */
    double coupling;        /* J = |K| >= 0                                    */
    uint64_t seed;          /* Philox key                                      */
    int32_t n_therm;        /* unmeasured sweeps                               */
    int32_t n_meas;         /* measured sweeps                                 */
    int32_t n_levels;       /* blockings per sample, <= log2(L) - 1            */
    int32_t sweeps_per_bin; /* measured sweeps per record bin (>= 1); the job
                               has ceil(n_meas / sweeps_per_bin) bins         */
    int32_t sampler;        /* MCRG_SAMPLER_* (0 = Metropolis); Swendsen-Wang
                               jobs run on the HBM schedule at any L          */
} mcrg_job_desc;

typedef struct {
    int32_t threads;           /* workgroup size 64/128/256/512; 0 = 256       */
    int32_t sweeps_per_launch; /* sweeps per kernel launch; 0 = 250           */
    int32_t max_group;         /* upper bound on lattices per workgroup; 0 = 256 */
    int32_t max_slot_lds;      /* LDS bytes a workgroup of small lattices may take
                                  (lattices per workgroup halve until it fits);
                                  0 = 32 KiB per 256 threads                    */
    int32_t min_team;          /* lanes per lattice at least (1..threads, power
                                  of two); 0 = 16                               */
    int32_t max_team;          /* lanes per lattice at most (power of two,
                                  >= min_team); 0 = threads                     */
    int32_t threads_large;     /* workgroup size for lattices of L = 512
                                  (64/128/256/512/1024); 0 = 1024 when there
                                  are at most as many of them as compute
                                  units, else 512                               */
    int32_t autotune;          /* nonzero: the first launches each try one
                                  candidate configuration (timed with events)
                                  and the fastest is kept if it beats the
                                  configuration given here by more than 4 %;
                                  results are the same under every
                                  configuration                                 */
    int32_t sub_chunks;        /* parts each workgroup slot's sweeps of one
                                  launch are split into (0..64); 0 = automatic:
                                  up to 4 parts of at least 16 sweeps when the
                                  slots are more than the compute units but
                                  fewer than twice the resident workgroups, so
                                  that no CU holding one slot more than the
                                  others sets the launch time; else 1.  Parts
                                  of a slot run in order, possibly in
                                  different workgroups; results are the same
                                  for every value                               */
} mcrg_batch_opts;

mcrg_status mcrg_batch_create(mcrg_ctx ctx, const mcrg_job_desc *jobs, int32_t n_jobs,
                              const mcrg_batch_opts *opts, mcrg_batch *out);
mcrg_status mcrg_batch_destroy(mcrg_batch b);
/* Enqueues launches that advance every unfinished job by up to max_sweeps
 * sweeps (asynchronous).  *remaining receives the largest number of sweeps any
 * job still has to run after them. */
mcrg_status mcrg_batch_run(mcrg_batch b, int32_t max_sweeps, int32_t *remaining);
mcrg_status mcrg_batch_sync(mcrg_batch b);
/* Records of job j: n = ceil(n_meas / sweeps_per_bin) * mcrg_record_size(n_levels). */
mcrg_status mcrg_batch_get_records(mcrg_batch b, int32_t job, int64_t *host, int64_t n);
/* Spins of job j as +-1 int8, row-major, n = L*L. */
mcrg_status mcrg_batch_get_spins(mcrg_batch b, int32_t job, int8_t *host, int64_t n);
mcrg_status mcrg_batch_get_state(mcrg_batch b, int32_t job, uint32_t *sweep, uint32_t *n_measured);
/* Configuration in use (after autotuning: the one it kept; autotune = 1 while
 * candidates are still being tried).  No reference counterpart. */
mcrg_status mcrg_batch_get_opts(mcrg_batch b, mcrg_batch_opts *out);

/* ------------------------------------------------------------------ serving */
/* Continuous batching of MCRG jobs (mcrg_job_desc) on one device.  Jobs are
 * submitted at any time from any thread; a scheduler admits queued jobs in
 * arrival order while their spins and records fit the device-memory budget and
 * their share of LDS (slot LDS / lattices per slot) fits the LDS budget,
 * advances every admitted job by
 * one time slice per launch of the LDS-resident schedule (jobs of different L
 * and coupling share each launch; a slot of G lattices of side L runs
 * min(sweeps_per_launch, max(1, launch_quantum / t_sweep)) sweeps, with the
 * slot's measured-sweep time modelled as 8 us + 75 ns x (sites each of its
 * lanes sweeps), so every workgroup works about launch_quantum), and retires
 * finished jobs
 * between launches: their records are copied to host memory and their device
 * memory is reused by the next admissions.  A job's results are bit-identical
 * to the same job run alone (its own seed and counters).  No reference
 * counterpart: the reference runs one calc_critical_exponent per process
 * (src/main.cpp:15-17). */
typedef struct mcrg_server_s *mcrg_server;

typedef struct {
    int64_t hbm_budget_bytes;  /* device memory for resident jobs; 0 = 4 GiB          */
    int64_t lds_budget_bytes;  /* summed LDS share of resident jobs; 0 = two full
                                  waves of workgroups (2 x CUs x LDS per CU)          */
    int64_t launch_quantum;    /* microseconds of work per workgroup and launch;
                                  0 = 4000                                            */
    int32_t max_jobs;          /* jobs resident at once; 0 = 65536                    */
    int32_t threads;           /* workgroup size; 0 = 256                             */
    int32_t sweeps_per_launch; /* most sweeps of one job per launch; 0 = 65536        */
    int32_t threads_large;     /* workgroup size for L >= 256 jobs; 0 = adaptive:
                                  1024 while no job waits for admission, 512
                                  while some do (saturated device)            */
} mcrg_server_opts;

enum { MCRG_JOB_QUEUED = 0, MCRG_JOB_RUNNING = 1, MCRG_JOB_DONE = 2, MCRG_JOB_CANCELLED = 3 };

typedef struct {
    int32_t state;         /* MCRG_JOB_*                                        */
    int32_t n_bins;        /* record bins of the job                            */
    double submit_s;       /* times in seconds since the server was created     */
    double start_s;        /* admission (0 while queued)                        */
    double finish_s;       /* records available (0 until done)                  */
} mcrg_job_status;

typedef struct {
    int64_t submitted, admitted, completed, cancelled, queued, running;
    int64_t launches;
    double site_updates;   /* completed sweeps x L^2 over all jobs              */
    double busy_s;         /* time with at least one job running                */
    double device_wait_s;  /* part of busy_s spent waiting for launches         */
    int64_t hbm_in_use;    /* bytes of the budget held by resident jobs         */
    int64_t lds_in_use;    /* LDS shares of the resident jobs                   */
} mcrg_server_stats;

mcrg_status mcrg_server_create(mcrg_ctx ctx, const mcrg_server_opts *opts, mcrg_server *out);
/* Stops the scheduler if running and frees everything. */
mcrg_status mcrg_server_destroy(mcrg_server s);
/* Queues a job; *job_id identifies it from now on.  MCRG_ERR_INVALID_ARG for a
 * malformed job, MCRG_ERR_OUT_OF_BUDGET for one that can never fit. */
mcrg_status mcrg_server_submit(mcrg_server s, const mcrg_job_desc *job, uint64_t *job_id);
mcrg_status mcrg_server_poll(mcrg_server s, uint64_t job_id, mcrg_job_status *out);
/* Records of a finished job, n = n_bins * mcrg_record_size(n_levels);
 * MCRG_ERR_NOT_READY until it is done. */
mcrg_status mcrg_server_records(mcrg_server s, uint64_t job_id, int64_t *host, int64_t n);
/* Drops a finished job's host results, or cancels a queued or running one. */
mcrg_status mcrg_server_release(mcrg_server s, uint64_t job_id);
/* Ids of jobs finished since the last call, in completion order (at most
 * max_ids; the rest stay for the next call).  *n receives the count. */
mcrg_status mcrg_server_completed(mcrg_server s, uint64_t *ids, int32_t max_ids, int32_t *n);
/* Runs the scheduler on a thread of its own until mcrg_server_stop. */
mcrg_status mcrg_server_start(mcrg_server s);
mcrg_status mcrg_server_stop(mcrg_server s);
/* One scheduling iteration on the calling thread (when not started): admit,
 * queue one chunk, then wait for the previous chunk and retire the jobs it
 * finished (a job completes in the iteration after its last chunk, while the
 * next chunk runs).  *active = jobs resident, finished ones not yet retired
 * included. */
mcrg_status mcrg_server_step(mcrg_server s, int32_t *active);
mcrg_status mcrg_server_stats_get(mcrg_server s, mcrg_server_stats *out);

/* ------------------------------------------------------------------ FP64 solve */
/* Host arithmetic on records; no device work.  The connected correlators are
 * formed exactly from the int64 sums in 128-bit integers, then in FP64.
 *
 * Linearised RG matrix between levels `level` and level + 1 from one record
 * (e.g. the sum of all bins), restricted to the first n_ops operators (2 = NN,
 * NNN as the reference; 3 adds the plaquette):
 *   dSb_dK  = <S_{l+1} S_l^T> - <S_{l+1}><S_l>^T
 *   dSb_dKb = <S_{l+1} S_{l+1}^T> - <S_{l+1}><S_{l+1}>^T
 *   T = dSb_dKb^-1 dSb_dK                      (src/mcrg.cpp:114-123)
 * T is row-major, n_ops x n_ops.  Single- and double-counted operators give the
 * same T up to a similarity transform, hence the same eigenvalues. */
mcrg_status mcrg_rg_matrix(const int64_t *record, int32_t n_levels, int32_t level, int32_t n_ops,
                           double *T);
/* Largest real part of the eigenvalues of a real n x n matrix, n <= 3 (the
 * reference's EigenSolver + max of real parts, src/mcrg.cpp:126-130). */
double mcrg_leading_eigenvalue(int32_t n, const double *T);

typedef struct {
    double lambda;     /* leading eigenvalue of T at this level           */
    double yt;         /* ln(lambda) / ln 2                               */
    double nu;         /* ln 2 / ln(lambda) = 1 / yt (src/mcrg.cpp:131)   */
    double yt_err;     /* delete-one-bin jackknife errors (0 for 1 bin)   */
    double nu_err;
    int32_t n_jackknife;
} mcrg_rg_result;

/* Exponent at `level` from n_bins per-bin records (n_bins * record_size). */
mcrg_status mcrg_rg_exponent(const int64_t *records, int32_t n_bins, int32_t n_levels, int32_t level,
                             int32_t n_ops, mcrg_rg_result *out);

typedef struct {
    double mean;         /* per-sample mean of the observable                    */
    double var;          /* per-sample variance                                  */
    double tau_int;      /* integrated autocorrelation time in samples (sweeps),
                            from the coarsest binning that keeps >= 32 bins      */
    double tau_err;      /* its statistical error, tau_int * sqrt(2 / (bins - 1)) */
    double tau_fine;     /* the same from the bins as given: close to tau_int
                            when the bins are much longer than tau_int          */
    int32_t bin_samples; /* samples per bin behind tau_int                      */
    int32_t n_bins;      /* bins behind tau_int                                 */
} mcrg_tau_result;

/* Integrated autocorrelation time by binning, from n_bins per-bin records of
 * equal sample count (a shorter last bin is left out): with B samples per bin,
 * 2 tau_int = B var(bin means) / var(samples), tau_int = 1/2 + sum_t>=1 rho(t).
 * op 0, 1, 2 = NN, NNN, P at `level`; op 3 = |M| at level 0.  The bins must be
 * long against tau_int for the y_t jackknife to hold; this checks that.  No
 * counterpart in the reference, which does not bin. */
mcrg_status mcrg_tau_int(const int64_t *records, int32_t n_bins, int32_t n_levels, int32_t level, int32_t op,
                         mcrg_tau_result *out);

/* One matching step of the critical-point locator, approx_critical_point
 * (src/mcrg.cpp:191-310): `rec_large` is the record of an L lattice, `rec_small`
 * that of an L/2 lattice at the same coupling; at `level` n the L lattice blocked
 * n+1 times is matched with the L/2 lattice blocked n times through the NN
 * operator.  *dK is the reference's step (double-counted NN in numerator and
 * covariances, src/mcrg.cpp:293-298): the reference's new estimate is K + dK in
 * its sign convention (K = -J), i.e. J_new = J - dK. */
mcrg_status mcrg_critical_point_step(const int64_t *rec_large, int32_t nl_large, const int64_t *rec_small,
                                     int32_t nl_small, int32_t level, double *dK);
/* The same step from n_bins per-bin records of each chain (summed in 128-bit
 * integers, so totals beyond 2^63 stay exact). */
mcrg_status mcrg_critical_point_step_bins(const int64_t *rec_large, int32_t n_bins_large, int32_t nl_large,
                                          const int64_t *rec_small, int32_t n_bins_small, int32_t nl_small,
                                          int32_t level, double *dK);

/* ------------------------------------------------------------------ timing */
/* Kernel-duration statistics measured with HIP events on the stream that
 * launched the kernel, enabled per context.  kind: 0 = half-sweep kernel,
 * 1 = level-0 measurement kernel, 2 = levels/accumulate kernel,
 * 3 = batched resident kernel, 4 = strip kernel (one launch per
 * mcrg_chain_sweep call on the strip schedule). */
mcrg_status mcrg_ctx_set_timing(mcrg_ctx ctx, int32_t enable);
mcrg_status mcrg_ctx_kernel_time(mcrg_ctx ctx, int32_t kind, int64_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
