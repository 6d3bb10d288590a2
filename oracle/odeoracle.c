/*
 * odeoracle.c -- CPU restatement of the OrdinaryDiffEq.jl v7.7.0 algorithms on
 * the ensemble hot path.  TEST INFRASTRUCTURE ONLY: it is the parity checker
 * for the GPU engine and the CPU baseline timed beside it.  No product code
 * links, imports or calls anything in this directory.
 *
 * Reference files restated (paths relative to the OrdinaryDiffEq.jl tree):
 *   lib/OrdinaryDiffEqTsit5/src/tsit_perform_step.jl:188-262   (Tsit5 step, FSAL init)
 *   lib/OrdinaryDiffEqTsit5/src/tsit_tableaus.jl:45-101,244-276 (tableau, interpolant)
 *   lib/OrdinaryDiffEqTsit5/src/interpolants.jl:24-93           (dense output)
 *   lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_perform_step.jl:561-847 (Rodas step)
 *   lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_tableaus.jl:23-73 (Rodas5P tableau)
 *   lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_interpolants.jl:172-206
 *   lib/OrdinaryDiffEqDifferentiation/src/derivative_utils.jl:181-227,535-595,663-685,805-935,1011-1041
 *   lib/OrdinaryDiffEqCore/src/initdt.jl:17-298, integrators/integrator_interface.jl:644-662
 *   lib/OrdinaryDiffEqCore/src/solve.jl:128-203,904-946
 *   lib/OrdinaryDiffEqCore/src/integrators/integrator_utils.jl:84-414,493-677,1027-1034,1199-1256,1290-1314
 *   lib/OrdinaryDiffEqCore/src/integrators/controllers.jl:244-293,744-843; alg_utils.jl:407-879
 *   lib/DiffEqBase/src/calculate_residuals.jl:9-14, common_defaults.jl:69-75,190-225,
 *   lib/DiffEqBase/src/check_error.jl:69-118, utils.jl:90-94
 *   lib/OrdinaryDiffEqCore/src/integrators/integrator_utils.jl:215-239,720-725,1081-1145,1325-1338
 *       and integrator_interface.jl:5-80,443-446                (callbacks, reeval_fsal, terminate!)
 *   lib/DiffEqBase/src/callbacks.jl:91-146,157-640,645-719     (Discrete / Continuous / Vector callbacks)
 *   lib/OrdinaryDiffEqTsit5/src/tsit_perform_step.jl:1-38, lib/OrdinaryDiffEqRosenbrock/src/
 *       stiff_addsteps.jl:140-224                               (the event re-step)
 *
 * Parity unpinned (no reference test or fixture fixes them; the reference
 * cannot run here): FastPower's Float32 log2/exp2 rounding; LinearSolve's LU
 * operation order; MuladdMacro's nesting beyond its documented example; the
 * continuous-callback root finder (the reference's ModAB interval solve from
 * BracketingNonlinearSolve is restated as bisection to the same adjacent-float
 * bracket, so the event float agrees but the number of condition calls,
 * stats.ncondition, does not) and the interp_points sample times
 * (tprev + (i - 1) h here, Julia's range within an ulp).
 *
 * Floating-point conventions of the restatement (shared with the GPU engine by
 * specification, not by code):
 *   - every `@muladd` site is an explicit fma(); MuladdMacro nests a sum of
 *     products left to right: a1*k1 + a2*k2 + a3*k3 -> fma(a3,k3,fma(a2,k2,a1*k1));
 *   - the fused Rosenbrock helpers (_madd_terms) start from the base value:
 *     fma(cN,kN,...fma(c1,k1,base));
 *   - ODE_DEFAULT_NORM accumulates sum += x*x as an fma chain, then sqrt(sum/n);
 *   - Julia min/max semantics (NaN-propagating) are restated as jl_min/jl_max;
 *   - compile with -ffp-contract=off so no other contraction happens.
 * Third-party arithmetic restated (not vendored in the reference tree):
 *   FastPower.jl 1.1.x `fastpower` (called at controllers.jl:815-816): the
 *   published fast log2 of Goldberg (rational fit, "line 8 of table 2"),
 *   followed by exp2 in Float32; here exp2 is evaluated in double (oo_det_exp)
 *   and rounded to Float32.  initdt's 10^x and log10 use the same oo_det_log /
 *   oo_det_exp, so the GPU engine can reproduce every bit.
 *   LinearSolve.jl default dense LU: textbook right-looking LU with partial
 *   pivoting (first maximum on ties), multipliers scaled by the reciprocal of
 *   the pivot, trailing update a -= l*u as fma; triangular solves in LAPACK's
 *   column order (dtrsv).
 */
#include "odeoracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Julia scalar semantics                                                    */
/* ------------------------------------------------------------------------ */

/* Base.eps(x::Float64): spacing to the next float above |x|. */
double oo_jl_eps(double x) {
    x = fabs(x);
    if (!isfinite(x)) return NAN;
    if (x < DBL_MIN) return 4.9406564584124654e-324;
    int e;
    frexp(x, &e);
    return ldexp(1.0, e - 53);
}

/* Base.min / Base.max for IEEE floats: NaN-propagating, -0.0 < 0.0. */
static inline double jl_min(double x, double y) {
    if (isnan(x) || isnan(y)) return NAN;
    if (x == y) return signbit(x) ? x : y;
    return x < y ? x : y;
}
static inline double jl_max(double x, double y) {
    if (isnan(x) || isnan(y)) return NAN;
    if (x == y) return signbit(x) ? y : x;
    return x > y ? x : y;
}
/* Base.FastMath.max_fast as used inside calculate_residuals (@fastmath). */
static inline double max_fast(double x, double y) { return y > x ? y : x; }

/* ------------------------------------------------------------------------ */
/* det_log and det_exp follow fdlibm's e_log.c and e_exp.c (their argument
 * reduction, polynomial constants and reconstruction), which carry this notice:
 *   Copyright (C) 1993 (e_log.c), 2004 (e_exp.c) by Sun Microsystems, Inc.
 *   All rights reserved.
 *   Developed at SunSoft, a Sun Microsystems, Inc. business.
 *   Permission to use, copy, modify, and distribute this software is freely
 *   granted, provided that this notice is preserved. */
/* Deterministic log / exp (fdlibm-style reduction + minimax polynomials).   */
/* The GPU engine evaluates the same algorithm, so initdt's 10^x / log10 and */
/* fastpower's exp2 agree bit for bit between the two.  Accuracy ~1 ulp.     */
/* ------------------------------------------------------------------------ */
static const double DL_Lg1 = 6.666666666666735130e-01, DL_Lg2 = 3.999999999940941908e-01,
                    DL_Lg3 = 2.857142874366239149e-01, DL_Lg4 = 2.222219843214978396e-01,
                    DL_Lg5 = 1.818357216161805012e-01, DL_Lg6 = 1.531383769920937332e-01,
                    DL_Lg7 = 1.479819860511658591e-01;
static const double DL_ln2_hi = 6.93147180369123816490e-01, DL_ln2_lo = 1.90821492927058770002e-10;
static const double DE_P1 = 1.66666666666666019037e-01, DE_P2 = -2.77777777770155933842e-03,
                    DE_P3 = 6.61375632143793436117e-05, DE_P4 = -1.65339022054652515390e-06,
                    DE_P5 = 4.13813679705723846039e-08, DE_invln2 = 1.44269504088896338700e+00;
static const double D_LN10 = 2.302585092994046, D_LN2 = 0.6931471805599453;

double oo_det_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e = 0;
    if (x < DBL_MIN) { x *= 18014398509481984.0; e = -54; }
    uint64_t b;
    memcpy(&b, &x, 8);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s, w = z * z;
    double t1 = w * (DL_Lg2 + w * (DL_Lg4 + w * DL_Lg6));
    double t2 = z * (DL_Lg1 + w * (DL_Lg3 + w * (DL_Lg5 + w * DL_Lg7)));
    double R = t2 + t1;
    double hfsq = 0.5 * f * f;
    double dk = (double)e;
    return dk * DL_ln2_hi - ((hfsq - (s * (hfsq + R) + dk * DL_ln2_lo)) - f);
}

double oo_det_exp(double x) {
    if (isnan(x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019411) return 0.0;
    double k = rint(x * DE_invln2);
    double hi = x - k * DL_ln2_hi, lo = k * DL_ln2_lo;
    double r = hi - lo;
    double t = r * r;
    double c = r - t * (DE_P1 + t * (DE_P2 + t * (DE_P3 + t * (DE_P4 + t * DE_P5))));
    double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
    return ldexp(y, (int)k);
}

/* ------------------------------------------------------------------------ */
/* FastPower.jl restatement (controllers.jl:815-816 call sites)              */
/* ------------------------------------------------------------------------ */

float oo_fastlog2(float x) {
    const float a = 0.338953f, b = 2.198599f, c = 1.523692f;
    uint32_t ux1i;
    memcpy(&ux1i, &x, 4);
    uint32_t ex = (ux1i & 0x7F800000u) >> 23;
    uint32_t greater = ux1i & 0x00400000u;
    float signif, fexp;
    if (greater != 0u) {
        uint32_t ux2i = (ux1i & 0x007FFFFFu) | 0x3f000000u;
        memcpy(&signif, &ux2i, 4);
        fexp = (float)ex - 126.0f;
        signif = signif - 1.0f;
    } else {
        uint32_t ux2i = (ux1i & 0x007FFFFFu) | 0x3f800000u;
        memcpy(&signif, &ux2i, 4);
        fexp = (float)ex - 127.0f;
        signif = signif - 1.0f;
    }
    float t1 = a * signif;
    float t2 = t1 + b;
    float t3 = signif * t2;
    float t4 = signif + c;
    float t5 = t3 / t4;
    return fexp + t5;
}

double oo_fastpower(double x, double y) {
    if (x == 0.0) return 0.0;
    if (isinf(x) && isinf(y)) return INFINITY;
    float lg = oo_fastlog2((float)x);
    float prod = (float)y * lg;
    float r = (float)oo_det_exp((double)prod * D_LN2);
    return (double)r;
}

/* ------------------------------------------------------------------------ */
/* Tableaus (tsit_tableaus.jl:45-101, 244-276; rosenbrock_tableaus.jl:23-73)  */
/* ------------------------------------------------------------------------ */

static const double TS_c1 = 0.161, TS_c2 = 0.327, TS_c3 = 0.9, TS_c4 = 0.9800255409045097;
static const double TS_a21 = 0.161;
static const double TS_a31 = -0.008480655492356989, TS_a32 = 0.335480655492357;
static const double TS_a41 = 2.8971530571054935, TS_a42 = -6.359448489975075, TS_a43 = 4.3622954328695815;
static const double TS_a51 = 5.325864828439257, TS_a52 = -11.748883564062828, TS_a53 = 7.4955393428898365,
                    TS_a54 = -0.09249506636175525;
static const double TS_a61 = 5.86145544294642, TS_a62 = -12.92096931784711, TS_a63 = 8.159367898576159,
                    TS_a64 = -0.071584973281401, TS_a65 = -0.028269050394068383;
static const double TS_a71 = 0.09646076681806523, TS_a72 = 0.01, TS_a73 = 0.4798896504144996,
                    TS_a74 = 1.379008574103742, TS_a75 = -3.290069515436081, TS_a76 = 2.324710524099774;
static const double TS_bt1 = -0.00178001105222577714, TS_bt2 = -0.0008164344596567469,
                    TS_bt3 = 0.007880878010261995, TS_bt4 = -0.1447110071732629,
                    TS_bt5 = 0.5823571654525552, TS_bt6 = -0.45808210592918697,
                    TS_bt7 = 0.015151515151515152;
/* Tsit5Interp r_ij (coefficient of Theta^j in b_i(Theta)). */
static const double TS_R[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
    {0.0, 1.5, -4.0, 2.5},
};

void oo_tsit5_tableau(double* c, double* a, double* b, double* btilde) {
    double cc[7] = {0.0, TS_c1, TS_c2, TS_c3, TS_c4, 1.0, 1.0};
    double aa[7][7] = {{0}};
    aa[1][0] = TS_a21;
    aa[2][0] = TS_a31; aa[2][1] = TS_a32;
    aa[3][0] = TS_a41; aa[3][1] = TS_a42; aa[3][2] = TS_a43;
    aa[4][0] = TS_a51; aa[4][1] = TS_a52; aa[4][2] = TS_a53; aa[4][3] = TS_a54;
    aa[5][0] = TS_a61; aa[5][1] = TS_a62; aa[5][2] = TS_a63; aa[5][3] = TS_a64; aa[5][4] = TS_a65;
    aa[6][0] = TS_a71; aa[6][1] = TS_a72; aa[6][2] = TS_a73; aa[6][3] = TS_a74; aa[6][4] = TS_a75;
    aa[6][5] = TS_a76;
    double bt[7] = {TS_bt1, TS_bt2, TS_bt3, TS_bt4, TS_bt5, TS_bt6, TS_bt7};
    for (int i = 0; i < 7; i++) {
        c[i] = cc[i];
        b[i] = aa[6][i];
        btilde[i] = bt[i];
        for (int j = 0; j < 7; j++) a[i * 7 + j] = aa[i][j];
    }
}

void oo_tsit5_interp_coeffs(double* r) {
    for (int i = 0; i < 7; i++)
        for (int j = 0; j < 4; j++) r[i * 4 + j] = TS_R[i][j];
}

static const double RP_A[8][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {3.0, 0, 0, 0, 0, 0, 0, 0},
    {2.849394379747939, 0.45842242204463923, 0, 0, 0, 0, 0, 0},
    {-6.954028509809101, 2.489845061869568, -10.358996098473584, 0, 0, 0, 0, 0},
    {2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641, 0, 0, 0, 0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 0, 0, 0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 1, 0, 0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 1, 1, 0},
};
static const double RP_C[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {-14.155112264123755, 0, 0, 0, 0, 0, 0},
    {-17.97296035885952, -2.859693295451294, 0, 0, 0, 0, 0},
    {147.12150275711716, -1.41221402718213, 71.68940251302358, 0, 0, 0, 0},
    {165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306, 0, 0, 0},
    {24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125,
     -0.6610691825249471, 0, 0},
    {30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783,
     -19.097331116725623, -28.087943162872662, 0},
    {37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047,
     -40.06618937091002, -54.66780262877968, -9.48861652309627},
};
static const double RP_c[8] = {0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716,
                               0.4288403609558664, 1, 1, 1};
static const double RP_d[8] = {0.21193756319429014, -0.42387512638858027, -0.3384627126235924,
                               1.8046452872882734, 2.325825639765069, 0, 0, 0};
static const double RP_H[3][8] = {
    {25.948786856663858, -2.5579724845846235, 10.433815404888879, -2.3679251022685204,
     0.524948541321073, 1.1241088310450404, 0.4272876194431874, -0.17202221070155493},
    {-9.91568850695171, -0.9689944594115154, 3.0438037242978453, -24.495224566215796,
     20.176138334709044, 15.98066361424651, -6.789040303419874, -6.710236069923372},
    {11.419903575922262, 2.8879645146136994, 72.92137995996029, 80.12511834622643,
     -52.072871366152654, -59.78993625266729, -0.15582684282751913, 4.883087185713722},
};
static const double RP_gamma = 0.21193756319429014;

void oo_rodas5p_tableau(double* A, double* C, double* c, double* d, double* H, double* b,
                        double* btilde, double* gamma) {
    for (int i = 0; i < 8; i++) {
        for (int j = 0; j < 8; j++) {
            A[i * 8 + j] = RP_A[i][j];
            C[i * 8 + j] = (j < 7) ? RP_C[i][j] : 0.0;
        }
        c[i] = RP_c[i];
        d[i] = RP_d[i];
        b[i] = (i < 7) ? RP_A[7][i] : 1.0;
        btilde[i] = (i == 7) ? 1.0 : 0.0;
    }
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 8; i++) H[j * 8 + i] = RP_H[j][i];
    *gamma = RP_gamma;
}

/* ------------------------------------------------------------------------ */
/* Norm and residuals (common_defaults.jl:69-75, calculate_residuals.jl:9-14) */
/* ------------------------------------------------------------------------ */

static double ode_norm(const double* x, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fma(x[i], x[i], s);
    return sqrt(s / (double)(n > 1 ? n : 1));
}

/* EEst = ODE_DEFAULT_NORM(utilde ./ (abstol .+ max.(|u0|,|u1|) .* reltol)) */
static double residual_norm(const double* ut, const double* u0, const double* u1, int n,
                            double abstol, double reltol) {
    double s = 0.0;
    for (int i = 0; i < n; i++) {
        double den = fma(max_fast(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
        double r = ut[i] / den;
        s = fma(r, r, s);
    }
    return sqrt(s / (double)(n > 1 ? n : 1));
}

/* ------------------------------------------------------------------------ */
/* Dense LU (LinearSolve default, restated) -- see header comment            */
/* ------------------------------------------------------------------------ */

static void lu_factor(int n, double* A, int* piv) {
    for (int k = 0; k < n; k++) {
        int p = k;
        double amax = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(A[i * n + k]);
            if (v > amax) { amax = v; p = i; }
        }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; j++) {
                double tmp = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = tmp;
            }
        double inv = 1.0 / A[k * n + k];
        for (int i = k + 1; i < n; i++) {
            double l = A[i * n + k] * inv;
            A[i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[i * n + j] = fma(-l, A[k * n + j], A[i * n + j]);
        }
    }
}

static void lu_solve(int n, const double* LU, const int* piv, double* b) {
    for (int k = 0; k < n; k++) {
        int p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int i = 1; i < n; i++) {
        double acc = b[i];
        for (int j = 0; j < i; j++) acc = fma(-LU[i * n + j], b[j], acc);
        b[i] = acc;
    }
    /* back substitution in column order (LAPACK dtrsv 'U','N','N'):
     * x_j = x_j / U_jj, then x_i -= U_ij x_j for i < j, j = n-1 .. 0 */
    for (int j = n - 1; j >= 0; j--) {
        b[j] = b[j] / LU[j * n + j];
        for (int i = 0; i < j; i++) b[i] = fma(-LU[i * n + j], b[j], b[i]);
    }
}

int oo_lu_solve(int n, double* A, int* piv, double* b) {
    lu_factor(n, A, piv);
    lu_solve(n, A, piv, b);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Integrator state                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, alg;
    oo_rhs_fn f; oo_jac_fn jac; oo_tgrad_fn tgrad; void* user;
    const double* p;
    const oo_opts* o;
    double reltol, abstol, dtmax, dtmin, t0, tf;
    double tdir;          /* integration direction sign(tf - t0) (+1 when tf == t0) */
    /* controller constants (alg_utils.jl:407-879, controllers.jl:744-803) */
    double qmin, qmax, qmax_first, gamma, beta1, beta2, qoldinit, qsteady_min, qsteady_max;
    /* integrator scalars */
    double t, dt, dtpropose, tprev, tstop_target, EEst, q11, errold, dtcache;
    int64_t iter, success_iter;
    int accept_step, next_step_tstop, isout;
    int reeval_fsal;      /* a callback modified u: the next accept re-evaluates the FSAL stage */
    int event_last;       /* integrator.event_last_time: 1 + the continuous callback of the last event, 0 none */
    unsigned sim_mask;    /* components of that event when it was a vector callback's (prev_simultaneous_events) */
    double last_event_error;
    double* ctmp;         /* interpolated state for continuous-callback conditions */
    oo_result* res;
    /* vectors */
    double *u, *uprev, *tmp, *atmp;
    double* k[8];          /* Tsit5: k1..k7;  Rodas: ks1..ks8 */
    double *du, *dT, *lt, *fsal, *J, *W, *dense[3];
    int* piv;
    /* saving */
    double* t_out; double* u_out; int max_out;
    double* saveat_int; int n_saveat, saveat_pos;
    /* tstops queue (initialize_tstops, solve.jl:988-1030): user times strictly
     * inside (t0, tf), then tf, stored multiplied by tdir and sorted ascending,
     * so the next one in the integration direction is first; [tst_pos, n_tst)
     * are still queued.  saveat_int likewise (initialize_saveat). */
    double* tst; int n_tst, tst_pos;
} integ_t;

static inline double first_tstop(const integ_t* I) { return I->tst[I->tst_pos]; }

static void save_point(integ_t* I, double t, const double* v) {
    int s = I->res->nsaved;
    if (s < I->max_out) {
        if (I->t_out) I->t_out[s] = t;
        if (I->u_out) memcpy(I->u_out + (size_t)s * I->n, v, sizeof(double) * I->n);
    }
    I->res->nsaved = s + 1;
}

static inline void F(integ_t* I, double* du, const double* u, double t) {
    I->f(du, u, I->p, t, I->user);
    I->res->nf++;
}

/* initdt.jl:17-298 (_ode_initdt_iip, no mass matrix, no noise), called through
 * integrator_interface.jl:644-662 with dtmax = min(opts.dtmax, tstop - t). */
static double initdt(integ_t* I) {
    const int n = I->n;
    const double t = I->t, tdir = I->tdir;
    /* _determine_initdt passes dtmax = tdir min(|dtmax|, |first_tstop - tdir t|),
     * which _ode_initdt_iip multiplies by tdir again */
    const double dtmax_tdir = jl_min(fabs(I->dtmax), fabs(first_tstop(I) - tdir * t));
    const double dtmin = nextafter(jl_max(I->dtmin, oo_jl_eps(t)), INFINITY);
    const double smalldt = jl_max(dtmin, 1e-6);
    if (n == 0) return tdir * jl_max(smalldt, dtmin);
    double* sk = I->atmp;      /* scratch */
    double* f0 = I->du;
    double* f1 = I->dT;
    double* u1 = I->tmp;
    double* tmp = I->lt;
    for (int i = 0; i < n; i++) sk[i] = fma(fabs(I->u[i]), I->reltol, I->abstol);
    I->f(f0, I->u, I->p, t, I->user);
    for (int i = 0; i < n; i++) tmp[i] = I->u[i] / sk[i];
    double d0 = ode_norm(tmp, n);
    for (int i = 0; i < n; i++) tmp[i] = f0[i] / sk[i] * 1.0;
    double d1 = ode_norm(tmp, n);
    if (isnan(d1)) return tdir * dtmin;
    double dt0 = ((d0 < 1e-5) | (d1 < 1e-5)) ? smalldt : (d0 / d1) / 100.0;
    dt0 = jl_min(dt0, dtmax_tdir);
    if (dt0 < 10.0 * DBL_EPSILON) return tdir * jl_max(smalldt, dtmin);
    const double dt0_tdir = tdir * dt0;
    for (int i = 0; i < n; i++) u1[i] = fma(dt0_tdir, f0[i], I->u[i]);
    I->f(f1, u1, I->p, t + dt0_tdir, I->user);
    int same = 1;
    for (int i = 0; i < n; i++) if (!(f0[i] == f1[i])) { same = 0; break; }
    if (same) return tdir * jl_max(dtmin, 100.0 * dt0);
    for (int i = 0; i < n; i++) tmp[i] = (f1[i] - f0[i]) / sk[i] * 1.0;
    double d2 = ode_norm(tmp, n) / dt0 * 1.0;
    double mx = jl_max(d1, d2);
    double dt1;
    if (mx < 1e-15) /* `<= 1//10^15` evaluated exactly */
        dt1 = jl_max(1e-6, dt0 * 0.001);
    else
        dt1 = oo_det_exp(-(2.0 + oo_det_log(mx) / D_LN10) / 5.0 * D_LN10);
    return tdir * jl_max(dtmin, jl_min(jl_min(100.0 * dt0, dt1), dtmax_tdir));
}

/* integrator_utils.jl:268-324 (adaptive and dtchangeable branches); the queue
 * is never empty inside the solve loop (tf is its last entry) */
static void modify_dt_for_tstops(integ_t* I) {
    const double tdir = I->tdir;
    const double tdir_t = tdir * I->t;
    const double tdir_tstop = first_tstop(I);
    double distance = fabs(tdir_tstop - tdir_t);
    double tol = 0.0;
    if (isfinite(tdir_tstop) && isfinite(I->t))
        tol = 100.0 * oo_jl_eps(jl_max(fabs(I->t), fabs(tdir_tstop)));
    if (!I->o->adaptive) {   /* dtchangeable branch: always try dtcache */
        I->next_step_tstop = !(fabs(I->dtcache) + tol < distance);
        if (I->next_step_tstop) I->tstop_target = tdir * tdir_tstop;
        I->dt = tdir * jl_min(fabs(I->dtcache), distance);
        return;
    }
    double original_dt = fabs(I->dt);
    I->dtpropose = tdir * original_dt;
    if (original_dt + tol < distance) {
        I->next_step_tstop = 0;
    } else {
        I->next_step_tstop = 1;
        I->tstop_target = tdir * tdir_tstop;
    }
    I->dt = tdir * jl_min(original_dt, distance);
}

/* handle_tstop! (integrator_utils.jl:1290-1314), dtchangeable: pop every queued
 * copy of t once t reaches the first tstop.  Stepping past a tstop is an error
 * in the reference; it cannot happen here (dt is clipped to the distance). */
static void handle_tstop(integ_t* I) {
    const double tdir_t = I->tdir * I->t;
    if (I->tst_pos < I->n_tst && tdir_t == first_tstop(I))
        while (I->tst_pos < I->n_tst && tdir_t == first_tstop(I)) I->tst_pos++;
}

/* integrator_utils.jl:1243-1256 with utils.jl:90-94 (dtmax carries tdir's
 * sign; backward, dt = min(dt, |dtmin|) leaves a negative dt as it is) */
static void fix_dt_at_bounds(integ_t* I) {
    I->dt = (I->tdir > 0) ? jl_min(I->dtmax, I->dt) : jl_max(I->dtmax, I->dt);
    double dtmin_t = fabs(jl_max(oo_jl_eps(I->t), I->dtmin));
    I->dt = (I->tdir > 0) ? jl_max(I->dt, dtmin_t) : jl_min(I->dt, dtmin_t);
}

/* controllers.jl:805-821 */
static double stepsize_controller(integ_t* I) {
    double qmax = (I->success_iter == 0) ? I->qmax_first : I->qmax;
    double q;
    if (I->EEst == 0.0) {
        q = 1.0 / qmax;
    } else {
        double q11 = oo_fastpower(I->EEst, I->beta1);
        q = q11 / oo_fastpower(I->errold, I->beta2);
        I->q11 = q11;
        double x = q / I->gamma, lo = 1.0 / qmax, hi = 1.0 / I->qmin;
        q = (x > hi) ? hi : ((x < lo) ? lo : x);
    }
    return q;
}

/* controllers.jl:823-836 */
static double step_accept_controller(integ_t* I, double q) {
    if (I->qsteady_min <= q && q <= I->qsteady_max) q = 1.0;
    I->errold = jl_max(I->EEst, I->qoldinit);
    return I->dt / q;
}

/* controllers.jl:838-843 */
static void step_reject_controller(integ_t* I) {
    I->dt = I->dt / jl_min(1.0 / I->qmin, I->q11 / I->gamma);
}

/* integrator_utils.jl:1199-1210 */
static void calc_dt_propose(integ_t* I, double dtnew) {
    double dp = I->tdir * jl_min(fabs(I->dtmax), fabs(dtnew));
    I->dtpropose = I->tdir * jl_max(fabs(dp), fabs(jl_max(oo_jl_eps(I->t), I->dtmin)));
}

/* check_error.jl:69-118 (adaptive, force_dtmin = false, default unstable_check) */
static int check_error(integ_t* I) {
    if (isnan(I->dt)) return OO_RC_DTNAN;
    if (I->iter > I->o->maxiters) return OO_RC_MAXITERS;
    if (I->o->adaptive) {
        if (fabs(I->dt) <= fabs(I->dtmin) && (!I->accept_step || I->t + I->dt < I->tdir * first_tstop(I)))
            return OO_RC_DTLESSTHANMIN;
        else if (!I->accept_step && fabs(I->dt) <= fabs(oo_jl_eps(I->t)))
            return OO_RC_UNSTABLE;
    }
    if (I->accept_step) {
        for (int i = 0; i < I->n; i++)
            if (!isfinite(I->u[i])) return OO_RC_UNSTABLE;
    }
    return OO_RC_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Tsit5 perform_step! (tsit_perform_step.jl:204-262, @muladd)                */
/* ------------------------------------------------------------------------ */

static void tsit5_step(integ_t* I) {
    const int n = I->n;
    const double t = I->t, dt = I->dt;
    double *k1 = I->k[0], *k2 = I->k[1], *k3 = I->k[2], *k4 = I->k[3], *k5 = I->k[4],
           *k6 = I->k[5], *k7 = I->k[6];
    double *tmp = I->tmp, *uprev = I->uprev, *u = I->u;
    const double a = dt * TS_a21;
    for (int i = 0; i < n; i++) tmp[i] = fma(a, k1[i], uprev[i]);
    F(I, k2, tmp, fma(TS_c1, dt, t));
    for (int i = 0; i < n; i++) {
        double acc = TS_a31 * k1[i];
        acc = fma(TS_a32, k2[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    F(I, k3, tmp, fma(TS_c2, dt, t));
    for (int i = 0; i < n; i++) {
        double acc = TS_a41 * k1[i];
        acc = fma(TS_a42, k2[i], acc);
        acc = fma(TS_a43, k3[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    F(I, k4, tmp, fma(TS_c3, dt, t));
    for (int i = 0; i < n; i++) {
        double acc = TS_a51 * k1[i];
        acc = fma(TS_a52, k2[i], acc);
        acc = fma(TS_a53, k3[i], acc);
        acc = fma(TS_a54, k4[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    F(I, k5, tmp, fma(TS_c4, dt, t));
    for (int i = 0; i < n; i++) {
        double acc = TS_a61 * k1[i];
        acc = fma(TS_a62, k2[i], acc);
        acc = fma(TS_a63, k3[i], acc);
        acc = fma(TS_a64, k4[i], acc);
        acc = fma(TS_a65, k5[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    F(I, k6, tmp, t + dt);
    for (int i = 0; i < n; i++) {
        double acc = TS_a71 * k1[i];
        acc = fma(TS_a72, k2[i], acc);
        acc = fma(TS_a73, k3[i], acc);
        acc = fma(TS_a74, k4[i], acc);
        acc = fma(TS_a75, k5[i], acc);
        acc = fma(TS_a76, k6[i], acc);
        u[i] = fma(dt, acc, uprev[i]);
    }
    F(I, k7, u, t + dt);
    double* ut = I->atmp;
    for (int i = 0; i < n; i++) {
        double acc = TS_bt1 * k1[i];
        acc = fma(TS_bt2, k2[i], acc);
        acc = fma(TS_bt3, k3[i], acc);
        acc = fma(TS_bt4, k4[i], acc);
        acc = fma(TS_bt5, k5[i], acc);
        acc = fma(TS_bt6, k6[i], acc);
        acc = fma(TS_bt7, k7[i], acc);
        ut[i] = dt * acc;
    }
    I->EEst = residual_norm(ut, uprev, u, n, I->abstol, I->reltol);
}

/* interpolants.jl:52-64 with tsit5pre0 (@evalpoly = Horner via muladd) */
static void tsit5_interp(integ_t* I, double theta, double dt, double* out) {
    const int n = I->n;
    double th2 = theta * theta;
    double b[7];
    b[0] = theta * fma(theta, fma(theta, fma(theta, TS_R[0][3], TS_R[0][2]), TS_R[0][1]), TS_R[0][0]);
    for (int s = 1; s < 7; s++)
        b[s] = th2 * fma(theta, fma(theta, TS_R[s][3], TS_R[s][2]), TS_R[s][1]);
    for (int i = 0; i < n; i++) {
        double acc = I->k[0][i] * b[0];
        for (int s = 1; s < 7; s++) acc = fma(I->k[s][i], b[s], acc);
        out[i] = fma(dt, acc, I->uprev[i]);
    }
}

/* _ode_addsteps! for Tsit5Cache (tsit_perform_step.jl:1-38, @muladd) with
 * always_calc_begin: the stages k2..k7 of the step from uprev at tprev with
 * the current dt (k1 kept), as reeval_internals_due_to_modification! runs it
 * after change_t_via_interpolation!.  The reference charges no f calls here. */
static void tsit5_addsteps(integ_t* I) {
    const int n = I->n;
    const double t = I->tprev, dt = I->dt;
    double *k1 = I->k[0], *k2 = I->k[1], *k3 = I->k[2], *k4 = I->k[3], *k5 = I->k[4],
           *k6 = I->k[5], *k7 = I->k[6];
    double *tmp = I->tmp, *uprev = I->uprev;
    for (int i = 0; i < n; i++) tmp[i] = fma(dt, TS_a21 * k1[i], uprev[i]);
    I->f(k2, tmp, I->p, fma(TS_c1, dt, t), I->user);
    for (int i = 0; i < n; i++) {
        double acc = TS_a31 * k1[i];
        acc = fma(TS_a32, k2[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    I->f(k3, tmp, I->p, fma(TS_c2, dt, t), I->user);
    for (int i = 0; i < n; i++) {
        double acc = TS_a41 * k1[i];
        acc = fma(TS_a42, k2[i], acc);
        acc = fma(TS_a43, k3[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    I->f(k4, tmp, I->p, fma(TS_c3, dt, t), I->user);
    for (int i = 0; i < n; i++) {
        double acc = TS_a51 * k1[i];
        acc = fma(TS_a52, k2[i], acc);
        acc = fma(TS_a53, k3[i], acc);
        acc = fma(TS_a54, k4[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    I->f(k5, tmp, I->p, fma(TS_c4, dt, t), I->user);
    for (int i = 0; i < n; i++) {
        double acc = TS_a61 * k1[i];
        acc = fma(TS_a62, k2[i], acc);
        acc = fma(TS_a63, k3[i], acc);
        acc = fma(TS_a64, k4[i], acc);
        acc = fma(TS_a65, k5[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    I->f(k6, tmp, I->p, t + dt, I->user);
    for (int i = 0; i < n; i++) {
        double acc = TS_a71 * k1[i];
        acc = fma(TS_a72, k2[i], acc);
        acc = fma(TS_a73, k3[i], acc);
        acc = fma(TS_a74, k4[i], acc);
        acc = fma(TS_a75, k5[i], acc);
        acc = fma(TS_a76, k6[i], acc);
        tmp[i] = fma(dt, acc, uprev[i]);
    }
    I->f(k7, tmp, I->p, t + dt, I->user);
}

/* ------------------------------------------------------------------------ */
/* Rodas5P perform_step! (rosenbrock_perform_step.jl:729-847)                 */
/* ------------------------------------------------------------------------ */

static void rodas_step(integ_t* I) {
    const int n = I->n;
    const double t = I->t, dt = I->dt;
    double *uprev = I->uprev, *u = I->u, *du = I->du, *dT = I->dT, *lt = I->lt;
    double dtC[8][7], dtd[8];
    const double idt = 1.0 / dt;
    for (int s = 0; s < 8; s++) {
        for (int j = 0; j < 7; j++) dtC[s][j] = RP_C[s][j] * idt;
        dtd[s] = dt * RP_d[s];
    }
    const double dtgamma = dt * RP_gamma;
    /* fsalfirst = f(uprev, t) */
    F(I, I->fsal, uprev, t);
    /* calc_rosenbrock_differentiation! -> calc_W! -> do_newJW
     * (derivative_utils.jl:535-595, 805-935, 1011-1041) */
    int new_jac;
    if (I->iter <= 1) new_jac = 1;
    else {
        int errorfail = I->EEst > 1.0;
        new_jac = !(I->iter > 1 && errorfail);
    }
    if (new_jac) {
        I->jac(I->J, uprev, I->p, t, I->user);
        I->res->njacs++;
    }
    /* jacobian2W! (derivative_utils.jl:663-685), mass matrix = I */
    const double invdtgamma = 1.0 / dtgamma;
    for (int i = 0; i < n * n; i++) I->W[i] = I->J[i];
    for (int i = 0; i < n; i++) I->W[i * n + i] = fma(-1.0, invdtgamma, I->J[i * n + i]);
    I->res->nw++;
    /* calc_tderivative! (derivative_utils.jl:181-227) */
    if (new_jac) {
        if (I->tgrad) I->tgrad(dT, uprev, I->p, t, I->user);
        else for (int i = 0; i < n; i++) dT[i] = 0.0;
        if (!I->o->has_tgrad) I->res->nf++;   /* AD derivative costs one f call */
    }
    for (int i = 0; i < n; i++) lt[i] = I->fsal[i] + dtd[0] * dT[i];
    /* dolinsolve with A = W (refactorize) */
    lu_factor(n, I->W, I->piv);
    for (int i = 0; i < n; i++) I->tmp[i] = lt[i];
    lu_solve(n, I->W, I->piv, I->tmp);
    for (int i = 0; i < n; i++) I->k[0][i] = -I->tmp[i];
    I->res->nsolve++;
    for (int s = 1; s < 8; s++) {
        /* _stage_accum! : u = fma chain from uprev */
        for (int i = 0; i < n; i++) {
            double acc = uprev[i];
            for (int j = 0; j < s; j++) acc = fma(RP_A[s][j], I->k[j][i], acc);
            u[i] = acc;
        }
        F(I, du, u, fma(RP_c[s], dt, t));
        /* _stage_rhs! */
        for (int i = 0; i < n; i++) {
            double acc = fma(dtd[s], dT[i], du[i]);
            for (int j = 0; j < s; j++) acc = fma(dtC[s][j], I->k[j][i], acc);
            I->tmp[i] = acc;
        }
        lu_solve(n, I->W, I->piv, I->tmp);
        for (int i = 0; i < n; i++) I->k[s][i] = -I->tmp[i];
        I->res->nsolve++;
    }
    /* u = uprev + sum b_i ks_i ; b = A[8,1:7] ++ [1] */
    for (int i = 0; i < n; i++) {
        double acc = uprev[i];
        for (int j = 0; j < 7; j++) acc = fma(RP_A[7][j], I->k[j][i], acc);
        acc = fma(1.0, I->k[7][i], acc);
        u[i] = acc;
    }
    /* err = sum btilde_i ks_i (btilde = e8), base zero */
    for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int j = 0; j < 7; j++) acc = fma(0.0, I->k[j][i], acc);
        acc = fma(1.0, I->k[7][i], acc);
        du[i] = acc;
    }
    I->EEst = residual_norm(du, uprev, u, n, I->abstol, I->reltol);
    /* dense rows (calck: saveat non-empty, or continuous callbacks, whose
     * event search interpolates every step) */
    if (I->n_saveat > 0 || I->o->nccb > 0) {
        for (int r = 0; r < 3; r++)
            for (int i = 0; i < n; i++) {
                double acc = 0.0;
                for (int j = 0; j < 8; j++) acc = fma(RP_H[r][j], I->k[j][i], acc);
                I->dense[r][i] = acc;
            }
    }
}

/* _ode_addsteps! for RosenbrockCache (stiff_addsteps.jl:140-224) with
 * always_calc_begin, after change_t_via_interpolation!: the stages of the step
 * from uprev at tprev with the new dt, reusing the step's J and dT and its
 * fsalfirst f(uprev, tprev); W refactored.  That function has no @muladd: only
 * linsolve_tmp = fsalfirst + dtgamma dT (an explicit @muladd) is fused, the
 * stage sums, t + c dt, the stage right-hand sides and the dense rows
 * H ks round the product and the sum separately; dtC = C ./ dt.  The f calls
 * and solves are not counted in the stats (they go through the cache). */
static void rodas_addsteps(integ_t* I) {
    const int n = I->n;
    const double t = I->tprev, dt = I->dt;
    double *uprev = I->uprev, *du = I->du, *dT = I->dT, *lt = I->lt, *st = I->atmp;
    double dtC[8][7], dtd[8];
    for (int s = 0; s < 8; s++) {
        for (int j = 0; j < 7; j++) dtC[s][j] = RP_C[s][j] / dt;
        dtd[s] = dt * RP_d[s];
    }
    const double dtgamma = dt * RP_gamma;
    I->f(I->fsal, uprev, I->p, t, I->user);
    for (int i = 0; i < n; i++) lt[i] = fma(dtgamma, dT[i], I->fsal[i]);
    const double invdtgamma = 1.0 / dtgamma;
    for (int i = 0; i < n * n; i++) I->W[i] = I->J[i];
    for (int i = 0; i < n; i++) I->W[i * n + i] = fma(-1.0, invdtgamma, I->J[i * n + i]);
    lu_factor(n, I->W, I->piv);
    for (int i = 0; i < n; i++) I->tmp[i] = lt[i];
    lu_solve(n, I->W, I->piv, I->tmp);
    for (int i = 0; i < n; i++) I->k[0][i] = -I->tmp[i];
    for (int s = 1; s < 8; s++) {
        for (int i = 0; i < n; i++) {
            double acc = uprev[i];
            for (int j = 0; j < s; j++) acc = acc + RP_A[s][j] * I->k[j][i];
            st[i] = acc;
        }
        I->f(du, st, I->p, t + RP_c[s] * dt, I->user);
        for (int i = 0; i < n; i++) {
            double acc = du[i] + dtd[s] * dT[i];
            for (int j = 0; j < s; j++) acc = acc + dtC[s][j] * I->k[j][i];
            I->tmp[i] = acc;
        }
        lu_solve(n, I->W, I->piv, I->tmp);
        for (int i = 0; i < n; i++) I->k[s][i] = -I->tmp[i];
    }
    for (int r = 0; r < 3; r++)
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int j = 0; j < 8; j++) acc = acc + RP_H[r][j] * I->k[j][i];
            I->dense[r][i] = acc;
        }
}

/* rosenbrock_interpolants.jl:172-206, interp_order = 3 */
static void rodas_interp(integ_t* I, double theta, double* out) {
    const double th1 = 1.0 - theta;
    for (int i = 0; i < I->n; i++) {
        double in1 = fma(theta, I->dense[2][i], I->dense[1][i]);
        double in2 = fma(theta, in1, I->dense[0][i]);
        double in3 = fma(th1, in2, I->u[i]);
        out[i] = fma(theta, in3, th1 * I->uprev[i]);
    }
}

/* integrator_utils.jl:340-414 (_savevalues!, save_idxs = nothing); returns
 * savedexactly.  force_save: a callback's save of the current state. */
static int savevalues(integ_t* I, int force_save) {
    const oo_opts* o = I->o;
    int savedexactly = 0;
    while (I->saveat_pos < I->n_saveat && I->saveat_int[I->saveat_pos] <= I->tdir * I->t) {
        double curt = I->tdir * I->saveat_int[I->saveat_pos++];
        if (curt != I->t) {
            double theta = (curt - I->tprev) / I->dt;
            if (I->alg == OO_TSIT5) tsit5_interp(I, theta, I->dt, I->tmp);
            else rodas_interp(I, theta, I->tmp);
            save_point(I, curt, I->tmp);
        } else {
            if (curt == I->tf && !o->save_end) continue;
            savedexactly = 1;
            save_point(I, I->t, I->u);
        }
    }
    int s = I->res->nsaved;
    int empty = (s == 0);
    double last = (!empty && s - 1 < I->max_out && I->t_out) ? I->t_out[s - 1] : NAN;
    if (force_save ||
        (o->save_everystep && (empty || ((I->t != last || I->dt == 0.0) && (o->save_end || I->t != I->tf))))) {
        savedexactly = 1;
        save_point(I, I->t, I->u);
    }
    return savedexactly;
}

/* Continuous callbacks (Tsit5).  Julia's sign(): -1, 0, 1, NaN for NaN. */
static double jl_sign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : (x == 0 ? 0.0 : x)); }

/* get_condition (lib/DiffEqBase/src/callbacks.jl:91-146): the state at abst is
 * u at t, uprev at tprev, else the step's dense output; every call counts in
 * stats.ncondition. */
static double ccb_condition(integ_t* I, int k, double abst) {
    const double* v;
    if (abst == I->t) {
        v = I->u;
    } else if (abst == I->tprev) {
        v = I->uprev;
    } else {
        const double th = (abst - I->tprev) / I->dt;
        if (I->alg == OO_TSIT5) tsit5_interp(I, th, I->dt, I->ctmp);
        else rodas_interp(I, th, I->ctmp);
        v = I->ctmp;
    }
    I->res->ncondition++;
    return I->o->ccb_cond(v, I->p, abst, k, I->o->ccb_user);
}

/* is_event_occurrence (callbacks.jl:525-530) */
static int ccb_is_event(const int* fl, double prev_sign, double next_sign) {
    return ((prev_sign < 0 && fl[OO_CCB_HAS_AFFECT]) || (prev_sign > 0 && fl[OO_CCB_HAS_AFFECT_NEG])) &&
           prev_sign * next_sign <= 0;
}

/* find_callback_time for a ContinuousCallback (callbacks.jl:364-412) with
 * nudge_tprev (414-426), check_event_occurrence (428-451) and find_root
 * (480-497).  The root is bracketed down to adjacent floats (the interval
 * solve with abstol = reltol = 0) by bisection; LeftRootFind keeps the float on
 * the bottom side, RightRootFind the other.  The interp_points checks are at
 * tprev + (i - 1) h, h = (t - tprev) / (interp_points - 1), and t itself last.
 * Returns 1 when the step holds an event. */
static int ccb_find_time(integ_t* I, int k, double* cb_t, double* bottom_sign, double* residual) {
    const int* fl = I->o->ccb_flags + OO_CCB_NFLAGS * k;
    const double* tl = I->o->ccb_tol + 2 * k;
    double bottom_t = I->tprev;
    double bc = ccb_condition(I, k, bottom_t);
    if (I->event_last == k + 1) {
        if (fabs(bc - I->last_event_error) <= tl[0]) bottom_t = I->tprev + I->dt * tl[1];
        bc = ccb_condition(I, k, bottom_t);
    }
    const double bs = jl_sign(bc);
    double top_t = I->t;
    double ts = jl_sign(ccb_condition(I, k, top_t));
    int ev = ccb_is_event(fl, bs, ts);
    const int npts = fl[OO_CCB_INTERP_POINTS];
    if (!ev && npts != 0) {
        const double h = (I->t - I->tprev) / (double)(npts - 1);
        for (int i = 2; i <= npts; i++) {
            top_t = (i == npts) ? I->t : fma((double)(i - 1), h, I->tprev);
            ts = jl_sign(ccb_condition(I, k, top_t));
            ev = ccb_is_event(fl, bs, ts);
            if (ev) break;
        }
    }
    *bottom_sign = bs;
    if (!ev) { *cb_t = I->t; *residual = 0.0; return 0; }
    if (fl[OO_CCB_ROOTFIND] == 2 || ts == 0.0) { *cb_t = top_t; *residual = 0.0; return 1; }
    double l = bottom_t, r = top_t;
    for (int it = 0; it < 2200; it++) {
        const double m = l + 0.5 * (r - l);
        if (m == l || m == r) break;
        const double sm = jl_sign(ccb_condition(I, k, m));
        if (sm == 0.0) { l = m; r = m; break; }
        if (sm == bs) l = m; else r = m;
    }
    *cb_t = fl[OO_CCB_ROOTFIND] == 1 ? r : l;
    *residual = ccb_condition(I, k, *cb_t);
    return 1;
}

/* get_condition for a VectorContinuousCallback: all len values at abst. */
static void vcb_condition(integ_t* I, int k, double abst, double* out) {
    const double* v;
    if (abst == I->t) {
        v = I->u;
    } else if (abst == I->tprev) {
        v = I->uprev;
    } else {
        const double th = (abst - I->tprev) / I->dt;
        if (I->alg == OO_TSIT5) tsit5_interp(I, th, I->dt, I->ctmp);
        else rodas_interp(I, th, I->ctmp);
        v = I->ctmp;
    }
    I->res->ncondition++;
    I->o->vcb_cond(out, v, I->p, abst, k, I->o->ccb_user);
}

/* findall_events! (callbacks.jl:499-521) on the values at top_t: event i when
 * the bottom sign is nonzero and bottom_sign * value <= 0.  Returns the mask. */
static unsigned vcb_events(int len, const double* bsign, const double* val) {
    unsigned m = 0;
    for (int i = 0; i < len; i++)
        if (bsign[i] != 0 && bsign[i] * val[i] <= 0) m |= 1u << i;
    return m;
}

/* find_callback_time for a VectorContinuousCallback (callbacks.jl:239-362):
 * after this callback's own last event the bottom signs of the components
 * that fired are re-read at the nudged tprev (nudge_tprev on the value
 * closest to zero among them); the interp_points checks run unless every
 * component has an event at t; each event component is rooted to adjacent
 * floats (from the nudged tprev when it fired last time), the earliest root
 * wins, and components rooted at that same time are simultaneous, with
 * direction -sign(bottom).  ev receives the directions. */
static int vcb_find_time(integ_t* I, int k, double* cb_t, double* ev, double* residual) {
    const int* fl = I->o->ccb_flags + OO_CCB_NFLAGS * k;
    const double* tl = I->o->ccb_tol + 2 * k;
    const int len = fl[OO_CCB_VLEN];
    double bc[OO_VCB_MAXLEN], bs[OO_VCB_MAXLEN], top[OO_VCB_MAXLEN], ts[OO_VCB_MAXLEN];
    const double bottom_t = I->tprev;
    vcb_condition(I, k, bottom_t, bc);
    for (int i = 0; i < len; i++) { bs[i] = jl_sign(bc[i]); ev[i] = 0.0; }
    double nudged_t = bottom_t;
    const unsigned prev = (I->event_last == k + 1) ? I->sim_mask : 0u;
    if (I->event_last == k + 1) {
        double minval = 0.0, minabs = INFINITY;
        for (int i = 0; i < len; i++)
            if ((prev >> i) & 1u) {
                if (fabs(bc[i]) < minabs) { minabs = fabs(bc[i]); minval = bc[i]; }
            }
        if (fabs(minval - I->last_event_error) <= tl[0]) nudged_t = I->tprev + I->dt * tl[1];
        double tmp[OO_VCB_MAXLEN];
        vcb_condition(I, k, nudged_t, tmp);
        for (int i = 0; i < len; i++)
            if ((prev >> i) & 1u) bs[i] = jl_sign(tmp[i]);
    }
    double top_t = I->t;
    vcb_condition(I, k, top_t, top);
    unsigned mask = vcb_events(len, bs, top);
    const int npts = fl[OO_CCB_INTERP_POINTS];
    const unsigned all = (len >= 32) ? 0xffffffffu : ((1u << len) - 1u);
    if (npts != 0 && mask != all) {
        const double h = (I->t - I->tprev) / (double)(npts - 1);
        for (int i = 2; i <= npts; i++) {
            top_t = (i == npts) ? I->t : fma((double)(i - 1), h, I->tprev);
            vcb_condition(I, k, top_t, top);
            mask = vcb_events(len, bs, top);
            if (mask) break;
        }
    }
    for (int i = 0; i < len; i++) ts[i] = jl_sign(top[i]);
    if (!mask) { *cb_t = I->t; *residual = 0.0; return 0; }
    if (fl[OO_CCB_ROOTFIND] == 2) {
        for (int i = 0; i < len; i++)
            if ((mask >> i) & 1u) ev[i] = -bs[i];
        *cb_t = top_t; *residual = 0.0;
        return 1;
    }
    double cbt = nextafter(top_t, I->tdir > 0 ? INFINITY : -INFINITY), res = 0.0;   /* rightfloat */
    double tmp[OO_VCB_MAXLEN];
    for (int idx = 0; idx < len; idx++) {
        if (!((mask >> idx) & 1u)) continue;
        double cbi;
        if (ts[idx] == 0.0) {
            cbi = top_t;
        } else {
            double l = ((prev >> idx) & 1u) ? nudged_t : bottom_t, r = top_t;
            for (int it = 0; it < 2200; it++) {
                const double m = l + 0.5 * (r - l);
                if (m == l || m == r) break;
                vcb_condition(I, k, m, tmp);
                const double sm = jl_sign(tmp[idx]);
                if (sm == 0.0) { l = m; r = m; break; }
                if (sm == bs[idx]) l = m; else r = m;
            }
            cbi = fl[OO_CCB_ROOTFIND] == 1 ? r : l;
        }
        if (I->tdir * cbi < I->tdir * cbt)
            for (int i = 0; i < len; i++) ev[i] = 0.0;
        if (I->tdir * cbi <= I->tdir * cbt) {
            cbt = cbi;
            vcb_condition(I, k, cbi, tmp);
            res = tmp[idx];
            ev[idx] = -bs[idx];
        }
    }
    *cb_t = cbt; *residual = res;
    return 1;
}

/* apply_callback! for a ContinuousCallback (callbacks.jl:557-640):
 * set_proposed_dt! (dtrelax = 1), change_t_via_interpolation!
 * (integrator_interface.jl:5-40: u from the dense output, t, dt = t - tprev,
 * then reeval_internals_due_to_modification! recomputes the stages), saves,
 * affect! on an upcrossing / affect_neg! on a downcrossing.  Returns 1 for
 * terminate!. */
static int ccb_apply(integ_t* I, int k, double cb_t, double bs, const double* ev, int* saved_in_cb) {
    const oo_opts* o = I->o;
    const int* fl = o->ccb_flags + OO_CCB_NFLAGS * k;
    if (o->adaptive) {
        const double dtp = I->tdir * jl_max(nextafter(I->dtmin, INFINITY), I->tdir * I->dt);
        I->dtpropose = dtp;
        I->dtcache = dtp;
    }
    if (cb_t != I->t) {
        const double th = (cb_t - I->tprev) / I->dt;
        if (I->alg == OO_TSIT5) tsit5_interp(I, th, I->dt, I->ctmp);
        else rodas_interp(I, th, I->ctmp);
        memcpy(I->u, I->ctmp, sizeof(double) * I->n);
        I->t = cb_t;
        I->dt = I->t - I->tprev;
        if (I->alg == OO_TSIT5) tsit5_addsteps(I);
        else rodas_addsteps(I);
        I->reeval_fsal = 1;
    }
    int savedexactly = savevalues(I, 0);
    *saved_in_cb = 1;
    if (fl[OO_CCB_SAVE0] && !savedexactly) savevalues(I, 1);
    int modified = 1, term = 0;
    if (fl[OO_CCB_VLEN] > 0) {
        /* VectorContinuousCallback: affect!(integrator, simultaneous_events) */
        if (fl[OO_CCB_HAS_AFFECT]) term = o->vcb_affect(I->u, I->p, I->t, k, ev, o->ccb_user) != 0;
        else modified = 0;
    } else if (bs < 0) {
        if (fl[OO_CCB_HAS_AFFECT]) term = o->ccb_affect(I->u, I->p, I->t, k, 0, o->ccb_user) != 0;
        else modified = 0;
    } else if (bs > 0) {
        if (fl[OO_CCB_HAS_AFFECT_NEG]) term = o->ccb_affect(I->u, I->p, I->t, k, 1, o->ccb_user) != 0;
        else modified = 0;
    }
    if (modified) {
        I->reeval_fsal = 1;
        if (fl[OO_CCB_SAVE1]) { savevalues(I, 1); *saved_in_cb = 1; }
    }
    return term;
}

/* handle_callbacks! (integrator_utils.jl:1081-1132) for a CallbackSet of
 * DiscreteCallbacks, each applied as apply_discrete_callback!
 * (lib/DiffEqBase/src/callbacks.jl:645-683).  An affect! leaves
 * derivative_discontinuity set, so reeval_internals_due_to_modification!
 * (integrator_interface.jl:54-80) sets reeval_fsal.  Returns 1 when an affect!
 * called terminate!. */
static int handle_callbacks(integ_t* I) {
    const oo_opts* o = I->o;
    int saved_in_cb = 0, term = 0;
    if (o->nccb > 0) {
        /* find_first_continuous_callback (callbacks.jl:157-237): the earliest
         * event time wins, the first callback on ties */
        int win = -1;
        double tmin = 0.0, bs_win = 0.0, res_win = 0.0;
        double ev_win[OO_VCB_MAXLEN] = {0}, ev[OO_VCB_MAXLEN];
        for (int k = 0; k < o->nccb; k++) {
            double cbt, bs = 0.0, rs;
            const int vlen = o->ccb_flags[OO_CCB_NFLAGS * k + OO_CCB_VLEN];
            const int occ = vlen > 0 ? vcb_find_time(I, k, &cbt, ev, &rs) : ccb_find_time(I, k, &cbt, &bs, &rs);
            if (occ && (win < 0 || I->tdir * cbt < I->tdir * tmin)) {
                win = k; tmin = cbt; bs_win = bs; res_win = rs;
                for (int i = 0; i < OO_VCB_MAXLEN; i++) ev_win[i] = i < vlen ? ev[i] : 0.0;
            }
        }
        if (win >= 0) {
            I->last_event_error = res_win;
            I->event_last = win + 1;
            I->sim_mask = 0;
            const int vlen = o->ccb_flags[OO_CCB_NFLAGS * win + OO_CCB_VLEN];
            for (int i = 0; i < vlen; i++)
                if (ev_win[i] != 0.0) I->sim_mask |= 1u << i;
            if (ccb_apply(I, win, tmin, bs_win, ev_win, &saved_in_cb)) term = 1;
        } else {
            I->event_last = 0;
        }
    }
    for (int k = 0; k < o->ndcb; k++) {
        if (o->dcb_cond(I->u, I->p, I->t, k, o->dcb_user)) {
            int savedexactly = savevalues(I, 0);
            saved_in_cb = 1;
            if (o->dcb_save[2 * k] && !savedexactly) savevalues(I, 1);
            if (o->dcb_affect(I->u, I->p, I->t, k, o->dcb_user)) term = 1;
            I->reeval_fsal = 1;
            if (o->dcb_save[2 * k + 1]) { savevalues(I, 1); saved_in_cb = 1; }
        }
        I->res->ncondition++;
    }
    if (!saved_in_cb) savevalues(I, 0);
    return term;
}

static int cmp_double(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

void oo_default_opts(oo_opts* o, int alg) {
    memset(o, 0, sizeof(*o));
    o->alg = alg;
    o->reltol = 1e-3;
    o->abstol = 1e-6;
    o->dtmax = NAN;
    o->dtmin = 0.0;
    o->dt = 0.0;
    o->maxiters = 1000000;
    o->save_start = 1;
    o->save_end = 1;
    o->save_everystep = 0;
    o->save_end_user = -1;
    o->adaptive = 1;
}

int oo_solve(int n, oo_rhs_fn f, oo_jac_fn jac, oo_tgrad_fn tgrad, void* user,
             const double* u0, const double* p, double t0, double tf,
             const oo_opts* opts, double* t_out, double* u_out, int max_out,
             double* u_final, oo_result* res) {
    integ_t S;
    integ_t* I = &S;
    memset(I, 0, sizeof(*I));
    memset(res, 0, sizeof(*res));
    I->n = n; I->alg = opts->alg; I->f = f; I->jac = jac; I->tgrad = tgrad; I->user = user;
    I->p = p; I->o = opts; I->res = res;
    I->reltol = opts->reltol; I->abstol = opts->abstol;
    I->t0 = t0; I->tf = tf;
    I->tdir = (tf < t0) ? -1.0 : 1.0;
    /* dtmax defaults to tf - t0; a positive user dtmax takes tdir's sign (solve.jl:401) */
    I->dtmax = (isnan(opts->dtmax) || opts->dtmax <= 0.0) ? (tf - t0) : I->tdir * opts->dtmax;
    I->dtmin = opts->dtmin;
    I->t_out = t_out; I->u_out = u_out; I->max_out = max_out;
    /* controller defaults: qmin 1//5, qmax 10, qmax_first_step 10000, gamma 9//10,
     * beta2 = 2//(5*5), beta1 = 7//(10*5), qoldinit 1//10^4,
     * qsteady = [1,1] (Tsit5) or [1, 6//5] (Rodas5P, adaptive implicit) */
    I->qmin = 0.2; I->qmax = 10.0; I->qmax_first = 10000.0; I->gamma = 0.9;
    I->beta2 = 0.08; I->beta1 = 0.14; I->qoldinit = 1e-4;
    I->qsteady_min = 1.0; I->qsteady_max = (I->alg == OO_RODAS5P) ? 1.2 : 1.0;

    size_t nn = (size_t)(n > 0 ? n : 1);
    double* mem = (double*)calloc(nn * 24 + nn * nn * 2 + 8, sizeof(double));
    int* piv = (int*)calloc(nn, sizeof(int));
    double* m = mem;
    I->u = m; m += nn; I->uprev = m; m += nn; I->tmp = m; m += nn; I->atmp = m; m += nn;
    for (int s = 0; s < 8; s++) { I->k[s] = m; m += nn; }
    I->du = m; m += nn; I->dT = m; m += nn; I->lt = m; m += nn; I->fsal = m; m += nn;
    for (int r = 0; r < 3; r++) { I->dense[r] = m; m += nn; }
    I->ctmp = m; m += nn;
    I->J = m; m += nn * nn; I->W = m; m += nn * nn;
    I->piv = piv;

    /* initialize_saveat (solve.jl:1103-1123): tdir s for the times with
     * tdir t0 < tdir s <= tdir tf, sorted */
    const double tdir = I->tdir;
    I->saveat_int = NULL; I->n_saveat = 0;
    if (opts->nsaveat > 0) {
        I->saveat_int = (double*)malloc(sizeof(double) * opts->nsaveat);
        for (int i = 0; i < opts->nsaveat; i++) {
            double s = tdir * opts->saveat[i];
            if (tdir * t0 < s && s <= tdir * tf) I->saveat_int[I->n_saveat++] = s;
        }
        qsort(I->saveat_int, I->n_saveat, sizeof(double), cmp_double);
    }

    /* initialize_tstops (solve.jl:1019-1030): tdir s strictly inside, then tdir tf */
    I->tst = (double*)malloc(sizeof(double) * (size_t)(opts->ntstops + 1));
    I->n_tst = 0; I->tst_pos = 0;
    for (int i = 0; i < opts->ntstops; i++) {
        double s = tdir * opts->tstops[i];
        if (tdir * t0 < s && s < tdir * tf) I->tst[I->n_tst++] = s;
    }
    qsort(I->tst, I->n_tst, sizeof(double), cmp_double);
    I->tst[I->n_tst++] = tdir * tf;

    memcpy(I->u, u0, sizeof(double) * n);
    memcpy(I->uprev, u0, sizeof(double) * n);
    I->t = t0; I->tprev = t0;
    I->EEst = 1.0; I->q11 = 1.0; I->errold = I->qoldinit;
    I->iter = 0; I->success_iter = 0; I->accept_step = 0; I->next_step_tstop = 0; I->isout = 0;
    I->tstop_target = 0.0;

    if (opts->save_start) save_point(I, t0, I->u);
    /* initialize! (tsit_perform_step.jl:188-202): FSAL pre-start */
    if (I->alg == OO_TSIT5) F(I, I->k[0], I->uprev, I->t);
    /* handle_dt! (solve.jl:968-985) */
    I->dtcache = opts->dt;
    if (opts->dt != 0.0) {   /* user dt (NaN included: check_error reports DtNaN) */
        /* a positive dt of a backward solve takes tdir's sign (solve.jl:968-985) */
        const double dt = (opts->adaptive && opts->dt > 0.0 && tdir < 0) ? tdir * opts->dt : opts->dt;
        I->dt = dt; I->dtpropose = dt;
    } else {
        I->dt = initdt(I);
        I->dtpropose = I->dt;
        res->nf += 2;
    }
    res->dt_init = I->dt;

    int rc = OO_RC_SUCCESS, terminated = 0;
    /* solve! (solve.jl:904-946): outer loop over the tstops queue, inner loop
     * of step attempts while t < first_tstop */
    while (I->tst_pos < I->n_tst && rc == OO_RC_SUCCESS) {
      const double first = first_tstop(I);
      while (tdir * I->t < first) {
        /* loopheader! (integrator_utils.jl:84-127) */
        if (I->iter > 0) {
            if (I->accept_step) {
                I->success_iter++;
                /* apply_step! (175-203) */
                memcpy(I->uprev, I->u, sizeof(double) * n);
                I->dt = I->dtpropose;
                /* update_fsal! (215-239): reset_fsal! (1325-1338) after a callback
                 * modified u, else the FSAL copy; Rodas5P is not FSAL */
                if (I->alg == OO_TSIT5) {
                    if (I->reeval_fsal) F(I, I->k[0], I->u, I->t);
                    else memcpy(I->k[0], I->k[6], sizeof(double) * n);
                }
                I->reeval_fsal = 0;   /* loopfooter_reset! (720-725) */
                modify_dt_for_tstops(I);
            } else if (I->isout) {
                I->dt = I->dt * I->qmin;   /* handle_step_rejection! (integrator_utils.jl:135-136) */
            } else {
                step_reject_controller(I);
            }
        }
        I->iter++;
        fix_dt_at_bounds(I);
        modify_dt_for_tstops(I);
        rc = check_error(I);
        if (rc != OO_RC_SUCCESS) break;
        const int should_exit = I->next_step_tstop;
        if (I->next_step_tstop && fabs(I->dt) < oo_jl_eps(fabs(I->t))) {
            I->accept_step = 1;   /* handle_tstop_step!: skip the tiny step */
        } else if (I->alg == OO_TSIT5) {
            tsit5_step(I);
        } else {
            rodas_step(I);
        }
        /* _loopfooter! (integrator_utils.jl:597-677) */
        double ttmp = I->t + I->dt;
        if (!opts->adaptive) {   /* non-adaptive branch: every step is accepted */
            res->naccept++;
            I->tprev = I->t;
            if (I->next_step_tstop) { I->next_step_tstop = 0; I->t = I->tstop_target; }
            else I->t = ttmp;
            I->accept_step = 1;
            I->dtpropose = I->dt;
            if (handle_callbacks(I)) { terminated = 1; I->tst_pos = I->n_tst; break; }
            if (should_exit) break;
            continue;
        }
        double q = stepsize_controller(I);
        I->isout = opts->isoutofdomain ? (opts->isoutofdomain(I->u, I->p, ttmp, opts->isout_user) != 0) : 0;
        I->accept_step = !I->isout && (I->EEst <= 1.0);
        if (I->accept_step) {
            res->naccept++;
            I->tprev = I->t;
            if (I->next_step_tstop) I->dt = I->dtpropose;
            if (I->next_step_tstop) { I->next_step_tstop = 0; I->t = I->tstop_target; }
            else I->t = ttmp;
            double dtnew = step_accept_controller(I, q);
            calc_dt_propose(I, dtnew);
            /* terminate! empties the tstops queue, which ends both solve! loops */
            if (handle_callbacks(I)) { terminated = 1; I->tst_pos = I->n_tst; break; }
        } else {
            res->nreject++;
        }
        if (should_exit) break;
      }
      if (rc != OO_RC_SUCCESS) break;
      handle_tstop(I);
    }
    if (rc == OO_RC_SUCCESS) {
        /* postamble! -> solution_endpoint_match_cur_integrator! (493-588), after
         * terminate! as well */
        if (opts->save_end) {
            int s = res->nsaved;
            int doit = (s == 0);
            if (!doit) {
                double last = (s - 1 < max_out && t_out) ? t_out[s - 1] : NAN;
                int in_cache = 0;
                for (int i = 0; i < opts->nsaveat; i++) if (opts->saveat[i] == I->t) in_cache = 1;
                doit = (last != I->t) &&
                       (opts->save_end_user == 1 || in_cache || I->t == tf || opts->nsaveat == 0);
            }
            if (doit) save_point(I, I->t, I->u);
        }
    }
    if (rc == OO_RC_SUCCESS && terminated) rc = OO_RC_TERMINATED;
    res->retcode = rc;
    res->t_final = I->t;
    res->dt_last = I->dt;
    if (u_final) memcpy(u_final, I->u, sizeof(double) * n);
    free(I->saveat_int);
    free(I->tst);
    free(mem);
    free(piv);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Built-in problem families                                                  */
/* ------------------------------------------------------------------------ */

typedef struct { int n; } builtin_user;

/* Lorenz: p = (sigma, rho, beta) */
static void lorenz_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)t; (void)user;
    double x = u[0], y = u[1], z = u[2];
    du[0] = p[0] * (y - x);
    du[1] = x * (p[1] - z) - y;
    du[2] = x * y - p[2] * z;
}
static void lorenz_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)t; (void)user;
    J[0] = -p[0]; J[1] = p[0]; J[2] = 0.0;
    J[3] = p[1] - u[2]; J[4] = -1.0; J[5] = -u[0];
    J[6] = u[1]; J[7] = u[0]; J[8] = -p[2];
}
/* Robertson: p = (k1, k2, k3) */
static void rober_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)t; (void)user;
    double y1 = u[0], y2 = u[1], y3 = u[2];
    du[0] = -p[0] * y1 + p[2] * y2 * y3;
    du[1] = p[0] * y1 - p[1] * (y2 * y2) - p[2] * y2 * y3;
    du[2] = p[1] * (y2 * y2);
}
static void rober_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)t; (void)user;
    double y2 = u[1], y3 = u[2];
    J[0] = -p[0]; J[1] = p[2] * y3; J[2] = p[2] * y2;
    J[3] = p[0]; J[4] = -2.0 * p[1] * y2 - p[2] * y3; J[5] = -p[2] * y2;
    J[6] = 0.0; J[7] = 2.0 * p[1] * y2; J[8] = 0.0;
}
/* Oscillator chain, n = 2m: u = (x_1..x_m, v_1..v_m), p = (k, c, alpha),
 * fixed ends x_0 = x_{m+1} = 0:
 *   x_i' = v_i ;  v_i' = -k (2 x_i - x_{i-1} - x_{i+1}) - c v_i - alpha x_i^3 */
static void chain_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)t;
    int n = ((builtin_user*)user)->n, m = n / 2;
    const double k = p[0], c = p[1], al = p[2];
    for (int i = 0; i < m; i++) du[i] = u[m + i];
    for (int i = 0; i < m; i++) {
        double x = u[i];
        double s = 2.0 * x;
        if (i > 0) s = s - u[i - 1];
        if (i < m - 1) s = s - u[i + 1];
        du[m + i] = -k * s - c * u[m + i] - al * (x * x * x);
    }
}
static void chain_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)t;
    int n = ((builtin_user*)user)->n, m = n / 2;
    const double k = p[0], c = p[1], al = p[2];
    memset(J, 0, sizeof(double) * n * n);
    for (int i = 0; i < m; i++) J[i * n + m + i] = 1.0;
    for (int i = 0; i < m; i++) {
        double x = u[i];
        int r = m + i;
        J[r * n + i] = -k * 2.0 - al * (3.0 * (x * x));
        if (i > 0) J[r * n + i - 1] = k;
        if (i < m - 1) J[r * n + i + 1] = k;
        J[r * n + m + i] = -c;
    }
}
/* linear u' = p0 * u (componentwise) */
static void linear_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)t;
    int n = ((builtin_user*)user)->n;
    for (int i = 0; i < n; i++) du[i] = p[0] * u[i];
}
static void linear_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)t; (void)u;
    int n = ((builtin_user*)user)->n;
    memset(J, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) J[i * n + i] = p[0];
}
/* blow-up u' = exp(u) (check_error.jl test problem) */
static void expb_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)t; (void)p;
    int n = ((builtin_user*)user)->n;
    for (int i = 0; i < n; i++) du[i] = exp(u[i]);
}
static void expb_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)t; (void)p;
    int n = ((builtin_user*)user)->n;
    memset(J, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) J[i * n + i] = exp(u[i]);
}

static void ball_f(double* du, const double* u, const double* p, double t, void* user) {
    (void)p; (void)t; (void)user;
    du[0] = u[1];
    du[1] = -9.81;
}
static void ball_jac(double* J, const double* u, const double* p, double t, void* user) {
    (void)u; (void)p; (void)t; (void)user;
    J[0] = 0.0; J[1] = 1.0; J[2] = 0.0; J[3] = 0.0;
}
static double ball_cond(const double* u, const double* p, double t, int k, void* user) {
    (void)p; (void)t; (void)k; (void)user;
    return u[0];
}
static int ball_affect(double* u, const double* p, double t, int k, int neg, void* user) {
    (void)t; (void)k; (void)user;
    if (neg) u[1] = (-p[0]) * u[1];
    return 0;
}
static const int ball_cb_flags[OO_CCB_NFLAGS] = {0, 1, 0, 0, 0, 10, 0};
static const double ball_cb_tol[2] = {10.0 * 2.220446049250313e-16, 0.01};

static int builtin_fns(int prob, oo_rhs_fn* f, oo_jac_fn* j) {
    switch (prob) {
        case OO_PROB_LORENZ: *f = lorenz_f; *j = lorenz_jac; return 3;
        case OO_PROB_ROBERTSON: *f = rober_f; *j = rober_jac; return 3;
        case OO_PROB_CHAIN: *f = chain_f; *j = chain_jac; return 3;
        case OO_PROB_LINEAR: *f = linear_f; *j = linear_jac; return 1;
        case OO_PROB_EXPBLOWUP: *f = expb_f; *j = expb_jac; return 0;
        case OO_PROB_BALL: *f = ball_f; *j = ball_jac; return 2;
        default: return -1;
    }
}

int oo_solve_builtin(int prob, int n, const double* u0, const double* p, double t0, double tf,
                     const oo_opts* opts, double* t_out, double* u_out, int max_out,
                     double* u_final, oo_result* res) {
    oo_rhs_fn f; oo_jac_fn j;
    if (builtin_fns(prob, &f, &j) < 0) return -1;
    builtin_user bu = {n};
    oo_opts o2 = *opts;
    o2.has_tgrad = 0;
    if (prob == OO_PROB_BALL && o2.nccb == 0) {
        o2.nccb = 1; o2.ccb_cond = ball_cond; o2.ccb_affect = ball_affect;
        o2.ccb_flags = ball_cb_flags; o2.ccb_tol = ball_cb_tol; o2.ccb_user = NULL;
    }
    return oo_solve(n, f, j, NULL, &bu, u0, p, t0, tf, &o2, t_out, u_out, max_out, u_final, res);
}

/* ------------------------------------------------------------------------ */
/* Threaded ensemble driver (EnsembleThreads shape: workers pull chunks)      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int prob; int64_t N; int nmax; const int* nvec;
    const double* u0; int np; const double* p; const double* t0; const double* tf;
    const oo_opts* opts; int64_t chunk;
    double* u_final; int32_t* retcode; int64_t* naccept; int64_t* nreject; int64_t* nf;
    int64_t next;
    pthread_mutex_t mu;
} ens_t;

static void* ens_worker(void* arg) {
    ens_t* E = (ens_t*)arg;
    double u0[256], pp[16], uf[256];
    oo_opts o = *E->opts;
    o.save_start = 0; o.save_everystep = 0; o.nsaveat = 0; o.saveat = NULL;
    for (;;) {
        pthread_mutex_lock(&E->mu);
        int64_t b = E->next;
        E->next += E->chunk;
        pthread_mutex_unlock(&E->mu);
        if (b >= E->N) break;
        int64_t e = b + E->chunk < E->N ? b + E->chunk : E->N;
        for (int64_t j = b; j < e; j++) {
            int n = E->nvec ? E->nvec[j] : E->nmax;
            for (int i = 0; i < n; i++) u0[i] = E->u0[(size_t)i * E->N + j];
            for (int k = 0; k < E->np; k++) pp[k] = E->p[(size_t)k * E->N + j];
            oo_result r;
            oo_solve_builtin(E->prob, n, u0, pp, E->t0[j], E->tf[j], &o, NULL, NULL, 0, uf, &r);
            if (E->u_final)
                for (int i = 0; i < E->nmax; i++)
                    E->u_final[(size_t)i * E->N + j] = (i < n) ? uf[i] : 0.0;
            if (E->retcode) E->retcode[j] = r.retcode;
            if (E->naccept) E->naccept[j] = r.naccept;
            if (E->nreject) E->nreject[j] = r.nreject;
            if (E->nf) E->nf[j] = r.nf;
        }
    }
    return NULL;
}

double oo_ensemble_builtin(int prob, int64_t N, int nmax, const int* nvec,
                           const double* u0, int np, const double* p,
                           const double* t0, const double* tf, const oo_opts* opts,
                           int threads, int64_t chunk,
                           double* u_final, int32_t* retcode, int64_t* naccept,
                           int64_t* nreject, int64_t* nf) {
    ens_t E;
    E.prob = prob; E.N = N; E.nmax = nmax; E.nvec = nvec; E.u0 = u0; E.np = np; E.p = p;
    E.t0 = t0; E.tf = tf; E.opts = opts; E.chunk = chunk > 0 ? chunk : 1;
    E.u_final = u_final; E.retcode = retcode; E.naccept = naccept; E.nreject = nreject; E.nf = nf;
    E.next = 0;
    pthread_mutex_init(&E.mu, NULL);
    if (threads < 1) threads = 1;
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, ens_worker, &E);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(th);
    pthread_mutex_destroy(&E.mu);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* ------------------------------------------------------------------------ */
/* Vern9: the high-accuracy FP64 reference (reltol 1e-12) used to state the   */
/* engines' final-state error.  Tableau: Verner's 9(8) "most efficient" pair  */
/* as held by lib/OrdinaryDiffEqVerner/src/verner_tableaus.jl:3608-3725 (the  */
/* Float64 literals); stage structure verner_rk_perform_step.jl:1274-1402     */
/* (16 RHS per step, u from stages 1,8..15, error from stages 1,8..16).  Step  */
/* control: a PI controller of order 9 (controllers.jl:244-293 shape) with    */
/* the textbook HNW initial step; the controller only decides how finely the  */
/* reference is resolved, not its value at convergence.                       */
/* ------------------------------------------------------------------------ */

static const double V9C[16] = {0.0, 0.03462, 0.09702435063878045, 0.14553652595817068, 0.561,
                               0.22900791159048503, 0.544992088409515, 0.645, 0.48375, 0.06757, 0.25,
                               0.6590650618730999, 0.8206, 0.9012, 1.0, 1.0};
/* A[s][j]: coefficient of k_{j+1} in stage s+1 (zero-based) */
static double V9A[16][16];
static double V9B[16], V9BT[16];
static int v9_ready = 0;

static void v9_init(void) {
    if (v9_ready) return;
    memset(V9A, 0, sizeof(V9A));
    memset(V9B, 0, sizeof(V9B));
    memset(V9BT, 0, sizeof(V9BT));
#define A_(s, j, v) V9A[(s) - 1][(j) - 1] = (v)
    A_(2, 1, 0.03462);
    A_(3, 1, -0.03893354388572875); A_(3, 2, 0.13595789452450918);
    A_(4, 1, 0.03638413148954267); A_(4, 3, 0.10915239446862801);
    A_(5, 1, 2.0257639143939694); A_(5, 3, -7.638023836496291); A_(5, 4, 6.173259922102322);
    A_(6, 1, 0.05112275589406061); A_(6, 4, 0.17708237945550218); A_(6, 5, 0.0008027762409222536);
    A_(7, 1, 0.13160063579752163); A_(7, 4, -0.2957276252669636); A_(7, 5, 0.08781378035642955);
    A_(7, 6, 0.6213052975225274);
    A_(8, 1, 0.07166666666666667); A_(8, 6, 0.33055335789153195); A_(8, 7, 0.2427799754418014);
    A_(9, 1, 0.071806640625); A_(9, 6, 0.3294380283228177); A_(9, 7, 0.1165190029271823);
    A_(9, 8, -0.034013671875);
    A_(10, 1, 0.04836757646340646); A_(10, 6, 0.03928989925676164); A_(10, 7, 0.10547409458903446);
    A_(10, 8, -0.021438652846483126); A_(10, 9, -0.10412291746271944);
    A_(11, 1, -0.026645614872014785); A_(11, 6, 0.03333333333333333); A_(11, 7, -0.1631072244872467);
    A_(11, 8, 0.03396081684127761); A_(11, 9, 0.1572319413814626); A_(11, 10, 0.21522674780318796);
    A_(12, 1, 0.03689009248708622); A_(12, 6, -0.1465181576725543); A_(12, 7, 0.2242577768172024);
    A_(12, 8, 0.02294405717066073); A_(12, 9, -0.0035850052905728597); A_(12, 10, 0.08669223316444385);
    A_(12, 11, 0.43838406519683376);
    A_(13, 1, -0.4866012215113341); A_(13, 6, -6.304602650282853); A_(13, 7, -0.2812456182894729);
    A_(13, 8, -2.679019236219849); A_(13, 9, 0.5188156639241577); A_(13, 10, 1.3653531876033418);
    A_(13, 11, 5.8850910885039465); A_(13, 12, 2.8028087862720628);
    A_(14, 1, 0.4185367457753472); A_(14, 6, 6.724547581906459); A_(14, 7, -0.42544428016461133);
    A_(14, 8, 3.3432791530012653); A_(14, 9, 0.6170816631175374); A_(14, 10, -0.9299661239399329);
    A_(14, 11, -6.099948804751011); A_(14, 12, -3.002206187889399); A_(14, 13, 0.2553202529443446);
    A_(15, 1, -0.7793740861228848); A_(15, 6, -13.937342538107776); A_(15, 7, 1.2520488533793563);
    A_(15, 8, -14.691500408016868); A_(15, 9, -0.494705058533141); A_(15, 10, 2.2429749091462368);
    A_(15, 11, 13.367893803828643); A_(15, 12, 14.396650486650687); A_(15, 13, -0.79758133317768);
    A_(15, 14, 0.4409353709534278);
    A_(16, 1, 2.0580513374668867); A_(16, 6, 22.357937727968032); A_(16, 7, 0.9094981099755646);
    A_(16, 8, 35.89110098240264); A_(16, 9, -3.442515027624454); A_(16, 10, -4.865481358036369);
    A_(16, 11, -18.909803813543427); A_(16, 12, -34.26354448030452); A_(16, 13, 1.2647565216956427);
#undef A_
    V9B[0] = 0.014611976858423152; V9B[7] = -0.3915211862331339; V9B[8] = 0.23109325002895065;
    V9B[9] = 0.12747667699928525; V9B[10] = 0.2246434176204158; V9B[11] = 0.5684352689748513;
    V9B[12] = 0.058258715572158275; V9B[13] = 0.13643174034822156; V9B[14] = 0.030570139830827976;
    V9BT[0] = -0.005357988290444578; V9BT[7] = -2.583020491182464; V9BT[8] = 0.14252253154686625;
    V9BT[9] = 0.013420653512688676; V9BT[10] = -0.02867296291409493; V9BT[11] = 2.624999655215792;
    V9BT[12] = -0.2825509643291537; V9BT[13] = 0.13643174034822156; V9BT[14] = 0.030570139830827976;
    V9BT[15] = -0.04834231373823958;
    v9_ready = 1;
}

static double v9_rms(const double* x, const double* sc, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) { double r = x[i] / sc[i]; s += r * r; }
    return sqrt(s / (n > 0 ? n : 1));
}

/* One Vern9 solve of a built-in problem from t0 to tf.  Returns 0 on success,
 * 1 when maxsteps was exceeded, 2 when dt underflowed; *nsteps = accepted steps. */
int oo_vern9_builtin(int prob, int n, const double* u0, const double* p, double t0, double tf,
                     double reltol, double abstol, int64_t maxsteps, double* u_final, int64_t* nsteps) {
    oo_rhs_fn f; oo_jac_fn jfn;
    if (builtin_fns(prob, &f, &jfn) < 0 || n <= 0 || n > 256) return -1;
    v9_init();
    builtin_user bu = {n};
    double u[256], uprev[256], tmp[256], ut[256], sc[256], k[16][256];
    for (int i = 0; i < n; i++) uprev[i] = u0[i];
    double t = t0;
    /* HNW initial step, order 9 */
    f(k[0], uprev, p, t, &bu);
    for (int i = 0; i < n; i++) sc[i] = abstol + reltol * fabs(uprev[i]);
    double d0 = v9_rms(uprev, sc, n), d1 = v9_rms(k[0], sc, n);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    if (h0 > tf - t0) h0 = tf - t0;
    for (int i = 0; i < n; i++) tmp[i] = uprev[i] + h0 * k[0][i];
    f(k[1], tmp, p, t + h0, &bu);
    for (int i = 0; i < n; i++) ut[i] = k[1][i] - k[0][i];
    double d2 = v9_rms(ut, sc, n) / h0;
    double mx = d1 > d2 ? d1 : d2;
    double h1 = mx <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / mx, 1.0 / 10.0);
    double dt = fmin(100.0 * h0, h1);
    if (dt > tf - t0) dt = tf - t0;
    const double beta1 = 7.0 / 90.0, beta2 = 2.0 / 45.0, qmin = 0.2, qmax = 10.0, gamma = 0.9;
    double qold = 1e-4;
    int64_t acc = 0, it = 0;
    while (t < tf) {
        if (++it > maxsteps) { for (int i = 0; i < n; i++) u_final[i] = uprev[i]; *nsteps = acc; return 1; }
        if (t + dt > tf) dt = tf - t;
        if (dt < 1e-14 * fmax(fabs(t), 1.0)) { for (int i = 0; i < n; i++) u_final[i] = uprev[i]; *nsteps = acc; return 2; }
        f(k[0], uprev, p, t, &bu);
        for (int s = 1; s < 16; s++) {
            for (int i = 0; i < n; i++) {
                double a = 0.0;
                for (int q = 0; q < s; q++) a += V9A[s][q] * k[q][i];
                tmp[i] = uprev[i] + dt * a;
            }
            f(k[s], tmp, p, t + V9C[s] * dt, &bu);
        }
        for (int i = 0; i < n; i++) {
            double b = 0.0, e = 0.0;
            for (int q = 0; q < 16; q++) { b += V9B[q] * k[q][i]; e += V9BT[q] * k[q][i]; }
            u[i] = uprev[i] + dt * b;
            ut[i] = dt * e;
            sc[i] = abstol + reltol * fmax(fabs(uprev[i]), fabs(u[i]));
        }
        double EEst = v9_rms(ut, sc, n);
        if (isnan(EEst)) EEst = 1e10;
        double q11 = pow(EEst, beta1);
        if (EEst <= 1.0) {
            double q = q11 / pow(qold, beta2) / gamma;
            q = fmax(1.0 / qmax, fmin(1.0 / qmin, q));
            t += dt;
            for (int i = 0; i < n; i++) uprev[i] = u[i];
            qold = fmax(EEst, 1e-4);
            acc++;
            dt = dt / q;
        } else {
            double q = fmin(1.0 / qmin, q11 / gamma);
            dt = dt / q;
        }
    }
    for (int i = 0; i < n; i++) u_final[i] = uprev[i];
    *nsteps = acc;
    return 0;
}

typedef struct {
    int prob; int64_t N; int nmax; const int* nvec; const double* u0; int np; const double* p;
    const double* t0; const double* tf; double reltol, abstol; int64_t maxsteps;
    double* u_final; int32_t* status; int64_t* nsteps; int64_t next; pthread_mutex_t mu;
} v9ens_t;

static void* v9_worker(void* arg) {
    v9ens_t* E = (v9ens_t*)arg;
    double u0[256], pp[16], uf[256];
    for (;;) {
        pthread_mutex_lock(&E->mu);
        int64_t j = E->next++;
        pthread_mutex_unlock(&E->mu);
        if (j >= E->N) break;
        int n = E->nvec ? E->nvec[j] : E->nmax;
        for (int i = 0; i < n; i++) u0[i] = E->u0[(size_t)i * E->N + j];
        for (int k = 0; k < E->np; k++) pp[k] = E->p[(size_t)k * E->N + j];
        int64_t ns = 0;
        int st = oo_vern9_builtin(E->prob, n, u0, pp, E->t0[j], E->tf[j], E->reltol, E->abstol, E->maxsteps, uf, &ns);
        for (int i = 0; i < E->nmax; i++) E->u_final[(size_t)i * E->N + j] = (i < n) ? uf[i] : 0.0;
        if (E->status) E->status[j] = st;
        if (E->nsteps) E->nsteps[j] = ns;
    }
    return NULL;
}

double oo_vern9_ensemble(int prob, int64_t N, int nmax, const int* nvec, const double* u0, int np, const double* p,
                         const double* t0, const double* tf, double reltol, double abstol, int64_t maxsteps,
                         int threads, double* u_final, int32_t* status, int64_t* nsteps) {
    v9_init();
    v9ens_t E;
    E.prob = prob; E.N = N; E.nmax = nmax; E.nvec = nvec; E.u0 = u0; E.np = np; E.p = p; E.t0 = t0; E.tf = tf;
    E.reltol = reltol; E.abstol = abstol; E.maxsteps = maxsteps; E.u_final = u_final; E.status = status;
    E.nsteps = nsteps; E.next = 0;
    pthread_mutex_init(&E.mu, NULL);
    if (threads < 1) threads = 1;
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, v9_worker, &E);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(th);
    pthread_mutex_destroy(&E.mu);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

void oo_vern9_tableau(double* c /*16*/, double* a /*16x16*/, double* b /*16*/, double* btilde /*16*/) {
    v9_init();
    memcpy(c, V9C, sizeof(V9C));
    memcpy(a, V9A, sizeof(V9A));
    memcpy(b, V9B, sizeof(V9B));
    memcpy(btilde, V9BT, sizeof(V9BT));
}
