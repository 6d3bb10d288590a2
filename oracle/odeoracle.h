/*
 * odeoracle.h -- CPU restatement of OrdinaryDiffEq.jl's Tsit5 / Rodas5P step
 * control, used ONLY as test infrastructure (the parity checker and the CPU
 * baseline timed beside the GPU engine).  Nothing in the product links this.
 *
 * Every function in odeoracle.c cites the reference file:line it restates.
 * Paths are relative to the OrdinaryDiffEq.jl v7.7.0 tree.
 */
#ifndef ODEORACLE_H
#define ODEORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OO_TSIT5 = 0, OO_RODAS5P = 1 };

/* SciMLBase.ReturnCode enum order (Default=0, Success=1, ...). */
enum {
    OO_RC_DEFAULT = 0, OO_RC_SUCCESS = 1, OO_RC_TERMINATED = 2, OO_RC_DTNAN = 3,
    OO_RC_MAXITERS = 4, OO_RC_MAXNUMSUB = 5, OO_RC_DTLESSTHANMIN = 6, OO_RC_UNSTABLE = 7
};

/* Built-in right-hand sides (the benchmark problem families). */
enum { OO_PROB_USER = 0, OO_PROB_LORENZ = 1, OO_PROB_ROBERTSON = 2, OO_PROB_CHAIN = 3,
       OO_PROB_LINEAR = 4, OO_PROB_EXPBLOWUP = 5,
       /* bouncing ball u = (height, velocity), u' = (v, -9.81), p = (restitution);
        * built in with its floor event: a ContinuousCallback on u[0] with
        * affect_neg! v <- -p[0] v, no saves (unless opts already carry callbacks) */
       OO_PROB_BALL = 6 };

typedef void (*oo_rhs_fn)(double* du, const double* u, const double* p, double t, void* user);
/* J is row-major n x n: J[i*n + j] = d f_i / d u_j */
typedef void (*oo_jac_fn)(double* J, const double* u, const double* p, double t, void* user);
typedef void (*oo_tgrad_fn)(double* dT, const double* u, const double* p, double t, void* user);
/* DiscreteCallback k of a CallbackSet (lib/DiffEqBase/src/callbacks.jl:645-683):
 * condition(u, t) after an accepted step; affect! rewrites u in place and
 * returns nonzero when it calls terminate! (integrator_interface.jl:443-446). */
typedef int (*oo_dcb_cond_fn)(const double* u, const double* p, double t, int k, void* user);
typedef int (*oo_dcb_affect_fn)(double* u, const double* p, double t, int k, void* user);
/* ContinuousCallback k (SciMLBase; handled by lib/DiffEqBase/src/callbacks.jl:
 * 239-640): the condition's value at (u, t); affect! (neg = 0, upcrossing) or
 * affect_neg! (neg = 1) rewrites u and returns nonzero for terminate!. */
typedef double (*oo_ccb_cond_fn)(const double* u, const double* p, double t, int k, void* user);
typedef int (*oo_ccb_affect_fn)(double* u, const double* p, double t, int k, int neg, void* user);
/* VectorContinuousCallback k: the len condition values into out; affect!
 * with the event directions ev[i] in {-1, 0, 1} (0: no event). */
typedef void (*oo_vcb_cond_fn)(double* out, const double* u, const double* p, double t, int k, void* user);
typedef int (*oo_vcb_affect_fn)(double* u, const double* p, double t, int k, const double* ev, void* user);
enum { OO_CCB_HAS_AFFECT = 0, OO_CCB_HAS_AFFECT_NEG = 1, OO_CCB_SAVE0 = 2, OO_CCB_SAVE1 = 3,
       OO_CCB_ROOTFIND = 4 /* 0 left, 1 right, 2 none */, OO_CCB_INTERP_POINTS = 5,
       OO_CCB_VLEN = 6 /* 0: ContinuousCallback, > 0: VectorContinuousCallback of that length */,
       OO_CCB_NFLAGS = 7 };
#define OO_VCB_MAXLEN 32

typedef struct {
    int alg;              /* OO_TSIT5 / OO_RODAS5P */
    double reltol, abstol;
    double dtmax;         /* <= 0 or NaN: tf - t0 (solve.jl:153) */
    double dtmin;         /* default 0 (solve.jl:152) */
    double dt;            /* 0: automatic initial dt (initdt.jl) */
    int64_t maxiters;     /* default 1_000_000 (solve.jl:161) */
    int save_start, save_end, save_everystep;
    int save_end_user;    /* -1: nothing (default rule), 0/1: user Bool */
    const double* saveat; /* sorted output times, may be NULL */
    int nsaveat;
    int has_tgrad;        /* Rodas stats: no nf increment for dT when 1 */
    int adaptive;         /* 0: fixed steps of opts.dt (solve.jl `adaptive = false`) */
    const double* tstops; /* user stopping times (any order, may be NULL); solve.jl:440-446 */
    int ntstops;
    /* isoutofdomain(u, p, t) (solve.jl kwarg; integrator_utils.jl:612-614): nonzero
     * rejects the step; the next dt is dt * qmin (integrator_utils.jl:135-136).  NULL: never. */
    int (*isoutofdomain)(const double* u, const double* p, double t, void* user);
    void* isout_user;
    int ndcb;                    /* discrete callbacks, applied in this order (0: none) */
    oo_dcb_cond_fn dcb_cond;
    oo_dcb_affect_fn dcb_affect;
    const int* dcb_save;         /* save_positions of callback k: dcb_save[2k], dcb_save[2k+1] */
    void* dcb_user;
    int nccb;                    /* continuous callbacks (Tsit5), checked before the discrete ones */
    oo_ccb_cond_fn ccb_cond;
    oo_ccb_affect_fn ccb_affect;
    const int* ccb_flags;        /* [OO_CCB_NFLAGS * k + OO_CCB_*] */
    const double* ccb_tol;       /* [2k]: abstol, [2k + 1]: repeat_nudge */
    void* ccb_user;
    oo_vcb_cond_fn vcb_cond;     /* the vector callbacks' conditions and affects */
    oo_vcb_affect_fn vcb_affect;
} oo_opts;

typedef struct {
    int retcode;
    int64_t nf, naccept, nreject, njacs, nw, nsolve;
    int nsaved;
    double t_final;       /* integrator t at exit */
    double dt_last;       /* integrator dt at exit */
    double dt_init;       /* dt after handle_dt! (automatic initial step) */
    int64_t ncondition;   /* callback condition calls (stats.ncondition) */
} oo_result;

void oo_default_opts(oo_opts* o, int alg);

/* Solve one IVP.  t_out: max_out slots, u_out: max_out * n (row per save).
 * u_final (n) receives the integrator state at exit. */
int oo_solve(int n, oo_rhs_fn f, oo_jac_fn jac, oo_tgrad_fn tgrad, void* user,
             const double* u0, const double* p, double t0, double tf,
             const oo_opts* opts, double* t_out, double* u_out, int max_out,
             double* u_final, oo_result* res);

/* Solve with a built-in RHS (prob = OO_PROB_*); np is implied by prob. */
int oo_solve_builtin(int prob, int n, const double* u0, const double* p,
                     double t0, double tf, const oo_opts* opts,
                     double* t_out, double* u_out, int max_out,
                     double* u_final, oo_result* res);

/* Ensemble of built-in problems over `threads` workers with dynamic chunks of
 * trajectory indices (the EnsembleThreads scheduling shape).  SoA inputs:
 * u0[i*N + j] (component i of trajectory j), p[k*N + j]; nvec[j] = state
 * dimension of trajectory j (NULL: all n).  Outputs: u_final SoA (nmax x N),
 * retcode/naccept/nreject/nf per trajectory (any may be NULL).
 * Returns wall seconds. */
double oo_ensemble_builtin(int prob, int64_t N, int nmax, const int* nvec,
                           const double* u0, int np, const double* p,
                           const double* t0, const double* tf, const oo_opts* opts,
                           int threads, int64_t chunk,
                           double* u_final, int32_t* retcode, int64_t* naccept,
                           int64_t* nreject, int64_t* nf);

/* Exposed pieces for known-answer tests. */
double oo_fastpower(double x, double y);
float oo_fastlog2(float x);
double oo_jl_eps(double x);
double oo_det_log(double x);
double oo_det_exp(double x);
/* Dense LU with partial pivoting + solve, in place on A (row-major n x n). */
int oo_lu_solve(int n, double* A, int* piv, double* b);
/* Tableau access for the order-condition tests. */
void oo_tsit5_tableau(double* c /*7*/, double* a /*7x7*/, double* b /*7*/, double* btilde /*7*/);
void oo_rodas5p_tableau(double* A /*8x8*/, double* C /*8x8*/, double* c /*8*/, double* d /*8*/,
                        double* H /*3x8*/, double* b /*8*/, double* btilde /*8*/, double* gamma);
void oo_tsit5_interp_coeffs(double* r /*7x4, r[i*4 + j] = r_{i+1, j+1}*/);

/* Vern9 high-accuracy reference (reltol ~1e-12) for built-in problems; the
 * ensemble form runs `threads` workers and returns wall seconds.  status per
 * trajectory: 0 ok, 1 maxsteps exceeded, 2 dt underflow. */
int oo_vern9_builtin(int prob, int n, const double* u0, const double* p, double t0, double tf,
                     double reltol, double abstol, int64_t maxsteps, double* u_final, int64_t* nsteps);
double oo_vern9_ensemble(int prob, int64_t N, int nmax, const int* nvec, const double* u0, int np, const double* p,
                         const double* t0, const double* tf, double reltol, double abstol, int64_t maxsteps,
                         int threads, double* u_final, int32_t* status, int64_t* nsteps);
void oo_vern9_tableau(double* c /*16*/, double* a /*16x16*/, double* b /*16*/, double* btilde /*16*/);

#ifdef __cplusplus
}
#endif
#endif
