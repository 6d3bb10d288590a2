/*
 * odemi.h -- C ABI of the MI355X ensemble ODE engine (libodemi.so).
 *
 * This is the drop-in boundary for the `EnsembleThreads` slot of
 * `solve(EnsembleProblem(prob; prob_func), Tsit5() | Rodas5P(), ensemblealg;
 *  trajectories, reltol, abstol, saveat, ...)`.  A host front-end (the Python
 * package `odemi`, or the Julia `EnsembleMI355X` module through `ccall`)
 * evaluates `prob_func` on the host, packs the trajectories structure-of-arrays,
 * traces `f(du,u,p,t)` / `f.jac` / `f.tgrad` into HIP source and calls:
 *
 *   odemi_create            <- SciMLBase `EnsembleAlgorithm` instance (EnsembleMI355X())
 *   odemi_compile_rhs       <- the `ODEFunction` (f, jac, tgrad); the reference
 *                              specialises the solver on f at
 *                              lib/DiffEqBase/src/solve.jl:918- (promote_f)
 *   odemi_solve_batch       <- SciMLBase `__solve(::EnsembleProblem, alg,
 *                              ::EnsembleThreads; trajectories, kw...)`, which runs
 *                              `solve(prob_i, alg; kw...)`
 *                              (lib/DiffEqBase/src/solve.jl:609-670 ->
 *                              lib/OrdinaryDiffEqCore/src/solve.jl:1-12,904-946)
 *                              once per trajectory
 *   retcode[j]              <- `sol.retcode` (SciMLBase.ReturnCode order;
 *                              lib/DiffEqBase/src/check_error.jl:69-118)
 *   stats[k*N + j]          <- `sol.stats` (DEStats: nf, naccept, nreject, njacs,
 *                              nw, nsolve, ncondition; lib/DiffEqBase/src/stats.jl:1-60)
 *   odemi_serve_*           <- no counterpart in the reference: a continuous-batching
 *                              request loop for many concurrent ensemble solves
 *
 * Conventions: status return 0 = OK, negative = error, message through
 * odemi_last_error(ctx).  A bad trajectory never aborts the process: it gets its
 * own retcode.  Arrays are structure-of-arrays over the N trajectories:
 * x[i*N + j] is component i of trajectory j.  All pointers in one call are
 * either host pointers or device pointers (args.device_ptrs).  Calls on one ctx
 * serialise on its stream; distinct ctx values may be used from distinct threads.
 */
#ifndef ODEMI_H
#define ODEMI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct odemi_ctx odemi_ctx;
typedef struct odemi_rhs odemi_rhs;

/* algorithms */
#define ODEMI_TSIT5 0
#define ODEMI_RODAS5P 1
/* engines */
#define ODEMI_ENGINE_PLAIN 0      /* FP64, one kernel launch per step attempt, state in HBM */
#define ODEMI_ENGINE_PERSISTENT 1 /* persistent kernel + work queue, state in registers/LDS */
/* precisions (persistent engine) */
#define ODEMI_FP64 0
#define ODEMI_MIXED 1             /* Tsit5: FP32 stages, FP64 state/t/error/controller;
                                     Rodas5P: FP32 LU + FP64 iterative refinement */
/* odemi_compile_rhs flags */
#define ODEMI_COMPILE_TSIT5 1
#define ODEMI_COMPILE_RODAS5P 2
#define ODEMI_HAS_USER_TGRAD 4    /* f.tgrad supplied: no nf charge for dT */
#define ODEMI_COMPILE_WG 8        /* also build the workgroup-per-trajectory Rodas5P kernels */

/* SciMLBase.ReturnCode values used per trajectory */
#define ODEMI_RC_DEFAULT 0
#define ODEMI_RC_SUCCESS 1
#define ODEMI_RC_TERMINATED 2     /* a discrete callback's affect called terminate! */
#define ODEMI_RC_DTNAN 3
#define ODEMI_RC_MAXITERS 4
#define ODEMI_RC_DTLESSTHANMIN 6
#define ODEMI_RC_UNSTABLE 7

typedef struct {
    int alg;                  /* ODEMI_TSIT5 / ODEMI_RODAS5P */
    int engine;               /* ODEMI_ENGINE_* (plain on a workgroup problem: one launch per
                                 attempt through the workgroup kernels, start/end saves only) */
    int precision;            /* ODEMI_FP64 / ODEMI_MIXED */
    int device_ptrs;          /* 1: every array below is a device pointer */
    int64_t N;                /* trajectories */
    int nmax;                 /* state dimension (max over trajectories) */
    int np;                   /* parameters per trajectory */
    const int32_t* nvec;      /* per-trajectory dimension for uneven batches, or NULL */
    const double* u0;         /* [nmax][N] */
    const double* p;          /* [np][N] */
    const double* t0;         /* [N] */
    const double* tf;         /* [N]; tf < t0 integrates backward (tdir = -1) */
    double reltol, abstol;    /* solve.jl defaults: 1e-3, 1e-6 */
    double dtmax;             /* <= 0: tf - t0 (solve.jl:153); > 0 takes each trajectory's tdir sign */
    double dtmin;             /* 0 (solve.jl:152) */
    double dt;                /* != 0: initial dt; 0: automatic (initdt.jl) */
    int64_t maxiters;         /* 1_000_000 (solve.jl:161) */
    const double* saveat;     /* sorted output times shared by all trajectories, or NULL */
    int nsaveat;
    const double* tstops;     /* sorted user stopping times shared by all trajectories, or NULL
                                 (solve.jl:440-446 `tstops`; each trajectory keeps those in (t0, tf)) */
    int ntstops;
    int save_start, save_end, save_everystep;
    int save_end_user;        /* -1: `save_end = nothing`, 0/1: user Bool */
    int max_saves;            /* capacity of t_out/u_out per trajectory */
    double* t_out;            /* [max_saves][N] */
    double* u_out;            /* [max_saves][nmax][N] */
    int32_t* nsaved;          /* [N] points saved (may exceed max_saves: extra dropped) */
    double* u_final;          /* [nmax][N] integrator state at exit (may be NULL) */
    double* t_final;          /* [N] (may be NULL) */
    int32_t* retcode;         /* [N] */
    int32_t* stats;           /* [7][N]: nf, naccept, nreject, njacs, nw, nsolve, ncondition (may be NULL) */
    int refine_sweeps;        /* ODEMI_MIXED Rodas5P: FP64 refinement sweeps per stage solve (default 1) */
    int block;                /* threads per workgroup (0: engine default) */
    int blocks_per_cu;        /* persistent kernels: resident workgroups per CU (0: library default; capped by occupancy) */
    int sort_by_cost;         /* 1: hand out trajectories longest-span first */
    void* stream;             /* hipStream_t to run on, or NULL for the ctx stream */
    double kernel_ms;         /* out: device time (HIP events) of the persistent kernel, or of all plain-engine launches */
    int64_t launches;         /* out: kernel launches issued */
} odemi_solve_args;

odemi_ctx* odemi_create(int device);
void odemi_destroy(odemi_ctx* ctx);
const char* odemi_last_error(const odemi_ctx* ctx);
/* compute units, free and total HBM bytes of the ctx device */
int odemi_device_info(odemi_ctx* ctx, int* cu_count, size_t* free_bytes, size_t* total_bytes);
/* Version / ABI tag: 4 (odemi_server_opts with max_saves / max_tstops,
 * odemi_request_opts with save and stopping times). */
int odemi_abi_version(void);

/* Compile a traced problem with hipRTC for gfx950.  `problem_src` defines
 *   template<typename T> __device__ void odemi_f(T* du, const T* u, const T* p, T t);
 *   template<typename T> __device__ void odemi_jac(T* J, const T* u, const T* p, T t); (row-major n x n)
 *   template<typename T> __device__ void odemi_tgrad(T* dT, const T* u, const T* p, T t);
 * and, for uneven batches, the workgroup forms odemi_f_wg / odemi_jac_wg.
 * Optional, lane kernels only (Tsit5; Rodas5P without ODEMI_COMPILE_WG):
 *   #define ODEMI_HAS_ISOUT 1 with
 *   template<typename T, typename Q> __device__ bool odemi_isout(const T* u, const Q* p, T t);
 *     <- the `isoutofdomain` solve option (integrator_utils.jl:612-614)
 * Optional, workgroup kernels (ODEMI_COMPILE_WG): #define ODEMI_HAS_ISOUT_WG 1 with
 *   template<typename T> __device__ bool odemi_isout_wg(const T* u, const T* p, T t, int n);
 *     <- the same option, one case per state size n of the batch
 *   #define ODEMI_NDCB k with template<int K> struct OdemiDcb, K = 0..k-1, holding
 *     static constexpr int save0, save1, term;
 *     template<typename T, typename Q> static __device__ bool cond(const T* u, const Q* p, T t);
 *     template<typename T, typename Q> static __device__ void affect(T* u, const Q* p, T t);
 *     <- a CallbackSet of k DiscreteCallbacks (condition, affect! on u, save_positions,
 *        terminate! when term; lib/DiffEqBase/src/callbacks.jl:645-683)
 *   #define ODEMI_NCCB k with template<int K> struct OdemiCcb holding
 *     static constexpr int vlen, has_affect, has_affect_neg, save0, save1, rootfind (0 left,
 *       1 right, 2 none), interp_points, term, term_neg;
 *     static constexpr double abstol, repeat_nudge;
 *     template<typename T, typename Q> static __device__ T cond(const T* u, const Q* p, T t);
 *     affect / affect_neg as above; with vlen > 0 (VectorContinuousCallback, ODEMI_VCB_MAXLEN
 *     the largest) instead vcond(T* out, const T* u, const Q* p, T t) and
 *     bool vaffect(T* u, const Q* p, T t, const double* ev) returning terminate!
 *     <- ContinuousCallbacks (event root finding; lib/DiffEqBase/src/callbacks.jl:157-640)
 * Optional, workgroup kernels: #define ODEMI_NDCB_WG k with template<int K> struct OdemiDcbWg
 *     holding save0, save1, term and
 *     template<typename T> static __device__ bool cond(const T* u, const T* p, T t, int n);
 *     template<typename T> static __device__ void affect(T* u, const T* p, T t, int n);
 *   #define ODEMI_NCCB_WG k with OdemiCcbWg<K> holding OdemiCcb's constants, a
 *     T cond(u, p, t, n), affect and affect_neg (u, p, t, n); with vlen > 0
 *     (ODEMI_VCB_MAXLEN_WG the largest) instead vcond(T* out, u, p, t, n) and
 *     bool vaffect(u, p, t, const double* ev, n) returning terminate!
 *     <- the same DiscreteCallbacks, ContinuousCallbacks and VectorContinuousCallbacks,
 *        one case per state size n of an uneven batch, run by one thread on the state in LDS
 * n/np are the (maximum) state and parameter counts.  Code objects are cached
 * on disk by source hash ($ODEMI_CACHE_DIR).  compile_seconds may be NULL. */
odemi_rhs* odemi_compile_rhs(odemi_ctx* ctx, const char* problem_src, int n, int np, int flags,
                             double* compile_seconds);
void odemi_rhs_destroy(odemi_rhs* rhs);

/* hipRTC-compile a problem into the on-disk code-object cache without touching a
 * GPU (used by the build step to pre-populate the cache).  Returns 0 on success;
 * path_out receives the cache file, log_out the compiler log on failure. */
int odemi_rtc_codeobject(const char* problem_src, int n, int np, int flags, char* path_out, int path_cap,
                         double* seconds, char* log_out, int log_cap);

/* The hipRTC version that compiles (and keys the cache of) code objects.
 * Returns 0, or -1 when hipRTC cannot report it.  No counterpart in the
 * reference (diagnostics: which compiler built the code that ran). */
int odemi_rtc_version(int* major, int* minor);

/* Solve N independent IVPs (one EnsembleProblem solve).  Synchronous. */
int odemi_solve_batch(odemi_ctx* ctx, odemi_rhs* rhs, odemi_solve_args* args);

/* odemi_solve_batch without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as odemi_solve_batch, with these limits:
 *   - args.device_ptrs must be 1 (the caller passes inputs and outputs on the device);
 *   - persistent engine only (ODEMI_ENGINE_PERSISTENT);
 *   - the solve's scratch (work-queue counters and order, start-up stage buffers,
 *     the mixed workgroup kernel's Jacobian blocks) must fit in 64 MiB, the
 *     library's fixed per-GPU scratch array; larger solves return -2 with a message;
 *   - asynchronous: returns once its launches are enqueued on args.stream (NULL:
 *     the ctx stream); outputs are valid after that stream synchronises;
 *     kernel_ms is not measured (0).
 * Every call on one GPU shares that GPU's scratch array (it keeps the last
 * call's contents; the call clears what must start at zero): calls that overlap
 * on one GPU, from any ctx or stream, race.  It calls no hipMalloc / hipFree. */
int odemi_solve_batch_noalloc(odemi_ctx* ctx, odemi_rhs* rhs, odemi_solve_args* args);

/* Fill args with the reference's solve defaults for `alg`. */
void odemi_default_args(odemi_solve_args* args, int alg);

/* ---------------------------------------------------------------------------
 * Serving (no counterpart in the reference): continuous batching of many
 * concurrent ensemble requests over a device-resident slot pool.  A worker
 * thread admits arriving requests into free slots (admission bounded by the
 * pool, i.e. by the HBM budget), keeps every trajectory's integrator state
 * (t, dt, dtpropose, u, uprev, FSAL stage / J, dT, q11, errold, counters) on
 * the device between ticks, advances all active trajectories together, and
 * copies a request's results to the caller's buffers when its last trajectory
 * finishes.  A trajectory's result does not depend on what else is in flight.
 * A request may carry its own tolerances, stopping times and (lane servers)
 * save times (odemi_submit_opts, within the server's max_tstops / max_saves);
 * the other options are the server's.
 * ------------------------------------------------------------------------- */
typedef struct odemi_server odemi_server;
typedef struct {
    int alg;                  /* ODEMI_TSIT5 / ODEMI_RODAS5P (lane kernels, or the workgroup
                                 kernels when the problem was compiled with ODEMI_COMPILE_WG) */
    int precision;            /* ODEMI_FP64 / ODEMI_MIXED */
    double reltol, abstol, dtmax, dtmin;
    int64_t maxiters;
    int64_t max_slots;        /* slot pool size (trajectories in flight) */
    size_t hbm_budget_bytes;  /* > 0: cap the pool so it fits this many bytes */
    int tick_attempts;        /* step attempts per trajectory per tick */
    int refine_sweeps;
    int max_saves;            /* saved points per trajectory a request may ask for
                                 (save times + start + end; 0: final state only) */
    int max_tstops;           /* stopping times per request (0: none) */
} odemi_server_opts;

void odemi_server_default_opts(odemi_server_opts* opts, int alg);
odemi_server* odemi_server_create(odemi_ctx* ctx, odemi_rhs* rhs, const odemi_server_opts* opts);
/* Stops the worker; unfinished requests end as failed (-4).  No other call may
 * be in progress on the server. */
void odemi_server_destroy(odemi_server* server);
/* Submit N trajectories (host SoA inputs, copied before return; u0 is [n][N]
 * with n the compiled dimension).  nvec [N] gives each trajectory's own
 * dimension for a workgroup (uneven-n Rodas5P) server, or NULL for n.  Outputs
 * are written at completion: u_final [n][N], retcode [N], stats [7][N] (stats
 * may be NULL).  Returns a request id (> 0) or a negative error. */
int64_t odemi_submit(odemi_server* server, int64_t N, const int32_t* nvec, const double* u0, const double* p,
                     const double* t0, const double* tf, double* u_final, int32_t* retcode, int32_t* stats);
/* Per-request options (odemi_submit_opts).  No counterpart in the reference
 * beyond the per-solve keywords it mirrors: reltol, abstol, saveat,
 * save_start, save_end and tstops of `solve(prob, alg; kw...)`
 * (lib/OrdinaryDiffEqCore/src/solve.jl:128-203), shared by the request's
 * trajectories.
 *   reltol, abstol: <= 0 takes the server's;
 *   saveat [nsaveat]: save times, sorted ascending; save_start, save_end and
 *     save_end_user as in odemi_solve_args (the request's points follow the
 *     same rules as a batch solve with those arguments); with t_out non-NULL
 *     the saved points are written to t_out [max_saves][N], u_out
 *     [max_saves][n][N] and nsaved [N] (points past max_saves are counted in
 *     nsaved, not written), max_saves being the server's, which nsaveat +
 *     save_start + save_end must not exceed;
 *   tstops [ntstops]: stopping times, sorted ascending, at most the server's
 *     max_tstops.
 * Save times need a lane (Tsit5 / Rodas5P, fixed n) server; stopping times
 * are taken by every server. */
typedef struct {
    double reltol, abstol;
    const double* saveat;
    int64_t nsaveat;
    int save_start, save_end, save_end_user;
    double* t_out;
    double* u_out;
    int32_t* nsaved;
    const double* tstops;
    int64_t ntstops;
} odemi_request_opts;
/* odemi_submit with the request's own options (ro may be NULL). */
int64_t odemi_submit_opts(odemi_server* server, int64_t N, const int32_t* nvec, const double* u0, const double* p,
                          const double* t0, const double* tf, const odemi_request_opts* ro, double* u_final,
                          int32_t* retcode, int32_t* stats);
/* 1 done, 0 pending or running, < 0 error */
int odemi_poll(odemi_server* server, int64_t request);
/* Block until done (timeout_s <= 0: no timeout).  1 done, 0 timed out, < 0 error. */
int odemi_wait(odemi_server* server, int64_t request, double timeout_s);
/* steady-clock seconds of submission, admission and completion */
int odemi_request_times(odemi_server* server, int64_t request, double* t_submit, double* t_admit, double* t_done);
int odemi_server_info(odemi_server* server, int64_t* slots, int64_t* ticks, int64_t* pending, int64_t* active,
                      int64_t* finished);
/* Forget a finished request (its id becomes unknown).  -2 while it is queued
 * or running.  Records of finished requests are otherwise kept for
 * odemi_request_times until the server is destroyed. */
int odemi_request_release(odemi_server* server, int64_t request);
const char* odemi_server_last_error(odemi_server* server);

#ifdef __cplusplus
}
#endif
#endif
