// odemi_device.h -- device side of the MI355X ensemble ODE engine.
//
// This header is compiled at run time by hipRTC together with a traced
// problem (odemi_f / odemi_jac / odemi_tgrad and the ODEMI_* size macros), so
// every kernel below is specialised for one right-hand side.  It restates the
// step-control semantics of OrdinaryDiffEq.jl v7.7.0:
//   Tsit5 step        lib/OrdinaryDiffEqTsit5/src/tsit_perform_step.jl:188-262
//   Tsit5 tableau     lib/OrdinaryDiffEqTsit5/src/tsit_tableaus.jl:45-101,244-276
//   Tsit5 dense       lib/OrdinaryDiffEqTsit5/src/interpolants.jl:24-93
//   Rodas5P step      lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_perform_step.jl:729-847
//   Rodas5P tableau   lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_tableaus.jl:23-73
//   Rodas dense       lib/OrdinaryDiffEqRosenbrock/src/rosenbrock_interpolants.jl:172-206
//   J/W/dT            lib/OrdinaryDiffEqDifferentiation/src/derivative_utils.jl:181-227,535-595,663-685
//   loop              lib/OrdinaryDiffEqCore/src/solve.jl:904-946,
//                     integrators/integrator_utils.jl:84-414,493-677,1027-1034,1199-1256
//   PI controller     lib/OrdinaryDiffEqCore/src/integrators/controllers.jl:244-293,805-843
//   initdt            lib/OrdinaryDiffEqCore/src/initdt.jl:17-298
//   residual, norm    lib/DiffEqBase/src/calculate_residuals.jl:9-14, common_defaults.jl:69-75
//   retcodes          lib/DiffEqBase/src/check_error.jl:69-118
// Arithmetic conventions: each `@muladd` site is an explicit fma, products are
// summed left to right (fma(a3,k3,fma(a2,k2,a1*k1))), fused Rosenbrock sums start
// from their base value, the RMS norm is an fma chain; the program is compiled
// with -ffp-contract=off so no other fusion happens.  FP64 kernels therefore
// reproduce the CPU restatement bit for bit.
#pragma once

#ifndef ODEMI_NP
#define ODEMI_NP 1
#endif
#ifndef ODEMI_HAS_TGRAD
#define ODEMI_HAS_TGRAD 0
#endif

namespace odemi {

typedef long long i64;
typedef unsigned long long u64;

enum { RC_DEFAULT = 0, RC_SUCCESS = 1, RC_TERMINATED = 2, RC_DTNAN = 3, RC_MAXITERS = 4,
       RC_DTLESSTHANMIN = 6, RC_UNSTABLE = 7 };
enum { ST_NF = 0, ST_NACCEPT = 1, ST_NREJECT = 2, ST_NJACS = 3, ST_NW = 4, ST_NSOLVE = 5,
       ST_NCONDITION = 6, ST_COUNT = 7 };
constexpr int kReqOpts = 8;   // ints per serving request-table entry (Params::req_opts)

// Kernel argument block (one struct, passed by value).  All arrays are SoA over
// the N trajectories of the batch: x[i * N + j] is component i of trajectory j.
struct Params {
    i64 N;
    int nmax;                 // state dimension (upper bound for uneven batches)
    int np;
    const int* nvec;          // per-trajectory dimension (uneven batches) or null
    const double* u0;         // [nmax][N]
    const double* p;          // [np][N]
    const double* t0;         // [N]
    const double* tf;         // [N]
    const int* order;         // work-queue order (permutation of 0..N-1) or null
    double reltol, abstol;
    double dtmax;             // <= 0: tf - t0 per trajectory
    double dtmin;
    double dt_user;           // > 0: fixed initial dt (no initdt)
    i64 maxiters;
    const double* saveat;     // sorted, shared by all trajectories
    int nsaveat;
    const double* tstops;     // user stopping times, sorted, shared by all trajectories
    int ntstops;
    int save_start, save_end, save_everystep, save_end_user;  // save_end_user: -1 nothing
    int max_saves;
    double* t_out;            // [max_saves][N]
    double* u_out;            // [max_saves][nmax][N]
    int* nsaved;              // [N]
    double* u_final;          // [nmax][N]
    double* t_final;          // [N]
    int* retcode;             // [N]
    int* stats;               // [ST_COUNT][N]
    u64* counters;            // [0] work queue head, [1] live count (plain mode), [2] finished
    // plain (one kernel per step attempt) mode: device-resident integrator state
    double* st_scalar;        // [16][N]
    double* st_vec;           // [8][nmax][N]
    int* st_int;              // [8][N]
    int refine_sweeps;        // mixed Rosenbrock: FP64 refinement sweeps per solve
    int tick_attempts;        // serving: step attempts per slot per tick
    // serving (continuous batching): slot pool bookkeeping
    i64 slot0, nslots;        // init kernel: slot range to initialise
    const int* slot_req;      // [N] request table index of each slot
    u64* req_done;            // [R] finished trajectories per request
    const double* req_saveat; // [R][sa_cap] each request's save times (lane servers) or null
    const double* req_tstops; // [R][ts_cap] each request's stopping times
    const int* req_opts;      // [R][kReqOpts] nsaveat, ntstops, save_start, save_end, save_end_user
    int sa_cap, ts_cap;
    float* wg_jbuf;           // workgroup Gauss-Jordan kernel: [grid][LD][LD] FP32 Jacobian scratch, LD = 16 ceil(nmax/16)
    // start-up values computed ahead of the persistent lane kernels (or null):
    double* pre_dt;           // [N] initial dt (initdt or the user's dt)
    double* pre_k1;           // [nmax][N] Tsit5 FSAL stage f(u0), exact in FP64
    double* rec_in;           // queue-order input records (see rec_in_words) or null
    double* rec_out;          // queue-order result records (see rec_out_words) or null
    const int* inv;           // queue slot of each trajectory, with the records
};

// ---------------------------------------------------------------------------
// Julia scalar semantics
// ---------------------------------------------------------------------------

// Base.eps(::Float64), as selects (no divergent branches in the step loop):
// biased exponent e = 0 -> smallest subnormal, 1..52 -> 2^(e-1) in subnormal
// bits, 53..2046 -> exponent field e - 52, 2047 (Inf/NaN) -> NaN.
__device__ __forceinline__ double jl_eps(double x) {
    const u64 b = (u64)__double_as_longlong(x) & 0x7fffffffffffffffULL;
    const int e = (int)(b >> 52);
    const int sh = (e > 0 && e <= 52) ? e - 1 : 0;
    u64 r = (e > 52) ? ((u64)(e - 52) << 52) : (1ULL << sh);
    r = (e == 0x7ff) ? 0x7ff8000000000000ULL : r;
    return __longlong_as_double((i64)r);
}
__device__ __forceinline__ double nextfloat_pos(double x) {  // nextfloat for x >= 0 finite
    return __longlong_as_double(__double_as_longlong(x) + 1);
}
// Base.min / Base.max (NaN-propagating, -0.0 < 0.0), as selects
__device__ __forceinline__ double jl_min(double x, double y) {
    const bool take_x = (x < y) || (x == y && signbit(x));
    const double r = take_x ? x : y;
    return __builtin_isunordered(x, y) ? __longlong_as_double(0x7ff8000000000000LL) : r;
}
__device__ __forceinline__ double jl_max(double x, double y) {
    const bool take_x = (x > y) || (x == y && !signbit(x));
    const double r = take_x ? x : y;
    return __builtin_isunordered(x, y) ? __longlong_as_double(0x7ff8000000000000LL) : r;
}
template <typename T>
__device__ __forceinline__ T max_fast(T x, T y) { return y > x ? y : x; }

// FP64 reciprocal, quotient and square root from the hardware estimates
// (v_rcp_f64, v_rsq_f64) with Newton-Raphson refinement: FP64-accurate to about
// one ulp but not always the correctly rounded result.  The mixed-precision
// lane engine's error norm and controller use them (FP64 range and precision
// without the IEEE division and square-root sequences); the FP64 engines keep
// IEEE / and sqrt for bit parity with the CPU restatement.  Positive normal
// operands (tolerance-scaled denominators, step-size ratios).
__device__ __forceinline__ double rcp_nr(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double div_nr(double a, double b) {
    const double r = rcp_nr(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double sqrt_nr(double x) {   // x >= 0
    const bool tiny = x < 0x1p-1000;                       // keep the estimate off subnormals
    const double xs = tiny ? x * 0x1p+1000 : x;
    const double y = __builtin_amdgcn_rsq(xs);
    double g = xs * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, xs), h, g);
    g = fma(fma(-g, g, xs), h, g);
    g = tiny ? g * 0x1p-500 : g;
    return (x > 0.0 && x < __builtin_huge_val()) ? g : x;   // 0, inf, NaN pass through
}

// det_log and det_exp follow fdlibm's e_log.c and e_exp.c (their argument
// reduction, polynomial constants and reconstruction), which carry this notice:
//   Copyright (C) 1993 (e_log.c), 2004 (e_exp.c) by Sun Microsystems, Inc.
//   All rights reserved.
//   Developed at SunSoft, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this software is freely
//   granted, provided that this notice is preserved.
// Deterministic log / exp (fdlibm-style argument reduction + minimax
// polynomials, ~1 ulp).  Used wherever the controller needs an elementary
// function, so the CPU restatement (same algorithm) agrees bit for bit.
__device__ __forceinline__ double det_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -__longlong_as_double(0x7ff0000000000000LL)
                                    : __longlong_as_double(0x7ff8000000000000LL);
    if (isinf(x)) return x;
    int e = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; e = -54; }
    u64 b = (u64)__double_as_longlong(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((i64)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                 Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                 Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                 Lg7 = 1.479819860511658591e-01;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s, w = z * z;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    double R = t2 + t1;
    double hfsq = 0.5 * f * f;
    double dk = (double)e;
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}
__device__ __forceinline__ double det_exp(double x) {
    if (isnan(x)) return x;
    if (x > 709.782712893384) return __longlong_as_double(0x7ff0000000000000LL);
    if (x < -745.1332191019411) return 0.0;
    const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                 P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                 P5 = 4.13813679705723846039e-08, invln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    double k = rint(x * invln2);
    double hi = x - k * ln2_hi, lo = k * ln2_lo;
    double r = hi - lo;
    double t = r * r;
    double c = r - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
    double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
    return ldexp(y, (int)k);
}
constexpr double kLN10 = 2.302585092994046, kLN2 = 0.6931471805599453;

// FastPower.jl 1.1.x `fastpower` (controllers.jl:815-816): Goldberg's rational
// fast log2 in Float32, then exp2 of the Float32 product evaluated in double
// (det_exp) and rounded once to Float32.
__device__ __forceinline__ float fastlog2f(float x) {
    unsigned ux1i = __float_as_uint(x);
    unsigned ex = (ux1i & 0x7F800000u) >> 23;
    float signif, fexp;
    if ((ux1i & 0x00400000u) != 0u) {
        signif = __uint_as_float((ux1i & 0x007FFFFFu) | 0x3f000000u);
        fexp = (float)ex - 126.0f;
    } else {
        signif = __uint_as_float((ux1i & 0x007FFFFFu) | 0x3f800000u);
        fexp = (float)ex - 127.0f;
    }
    signif = signif - 1.0f;
    float t1 = 0.338953f * signif;
    float t2 = t1 + 2.198599f;
    float t3 = signif * t2;
    float t4 = signif + 1.523692f;
    float t5 = t3 / t4;
    return fexp + t5;
}
__device__ __forceinline__ double fastpower(double x, double y) {
    if (x == 0.0) return 0.0;
    if (isinf(x) && isinf(y)) return __longlong_as_double(0x7ff0000000000000LL);
    float lg = fastlog2f((float)x);
    float prod = (float)y * lg;
    return (double)(float)det_exp((double)prod * kLN2);
}
// The same fastpower with the hardware reciprocal in Goldberg's rational and
// the hardware exp2 (v_rcp_f32 / v_exp_f32, ~1 ulp): the mixed-precision lane
// engines' controller.  FastPower evaluates exp2 under @fastmath; this form is
// not bit-matched to the CPU restatement (those engines are checked to
// tolerance).
__device__ __forceinline__ float fastlog2_hw(double x) {
    const float xf = (float)x;
    const unsigned ux = __float_as_uint(xf);
    const unsigned ex = (ux & 0x7F800000u) >> 23;
    const bool hi = (ux & 0x00400000u) != 0u;
    const float signif = __uint_as_float((ux & 0x007FFFFFu) | (hi ? 0x3f000000u : 0x3f800000u)) - 1.0f;
    const float fexp = (float)ex - (hi ? 126.0f : 127.0f);
    const float t3 = signif * (0.338953f * signif + 2.198599f);
    return fexp + t3 * __builtin_amdgcn_rcpf(signif + 1.523692f);
}
__device__ __forceinline__ float fastpower_hw(double x, double y) {
    if (x == 0.0) return 0.0f;
    if (isinf(x) && isinf(y)) return __builtin_huge_valf();
    return __builtin_amdgcn_exp2f((float)y * fastlog2_hw(x));
}

// ---------------------------------------------------------------------------
// Tableaus
// ---------------------------------------------------------------------------
struct Tsit5Tab {
    static constexpr double c1 = 0.161, c2 = 0.327, c3 = 0.9, c4 = 0.9800255409045097;
    static constexpr double a21 = 0.161;
    static constexpr double a31 = -0.008480655492356989, a32 = 0.335480655492357;
    static constexpr double a41 = 2.8971530571054935, a42 = -6.359448489975075,
                            a43 = 4.3622954328695815;
    static constexpr double a51 = 5.325864828439257, a52 = -11.748883564062828,
                            a53 = 7.4955393428898365, a54 = -0.09249506636175525;
    static constexpr double a61 = 5.86145544294642, a62 = -12.92096931784711,
                            a63 = 8.159367898576159, a64 = -0.071584973281401,
                            a65 = -0.028269050394068383;
    static constexpr double a71 = 0.09646076681806523, a72 = 0.01, a73 = 0.4798896504144996,
                            a74 = 1.379008574103742, a75 = -3.290069515436081,
                            a76 = 2.324710524099774;
    static constexpr double bt1 = -0.00178001105222577714, bt2 = -0.0008164344596567469,
                            bt3 = 0.007880878010261995, bt4 = -0.1447110071732629,
                            bt5 = 0.5823571654525552, bt6 = -0.45808210592918697,
                            bt7 = 0.015151515151515152;
};

__constant__ double kTsitR[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
    {0.0, 1.5, -4.0, 2.5},
};

// Rodas5P (Steinebach 2023), transformed Hairer-Wanner form.
struct RodasTab {
    static constexpr double gamma = 0.21193756319429014;
};
__constant__ double kRA[8][8] = {
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
__constant__ double kRC[8][7] = {
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
__constant__ double kRc[8] = {0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716,
                              0.4288403609558664, 1, 1, 1};
__constant__ double kRd[8] = {0.21193756319429014, -0.42387512638858027, -0.3384627126235924,
                              1.8046452872882734, 2.325825639765069, 0, 0, 0};
__constant__ double kRH[3][8] = {
    {25.948786856663858, -2.5579724845846235, 10.433815404888879, -2.3679251022685204,
     0.524948541321073, 1.1241088310450404, 0.4272876194431874, -0.17202221070155493},
    {-9.91568850695171, -0.9689944594115154, 3.0438037242978453, -24.495224566215796,
     20.176138334709044, 15.98066361424651, -6.789040303419874, -6.710236069923372},
    {11.419903575922262, 2.8879645146136994, 72.92137995996029, 80.12511834622643,
     -52.072871366152654, -59.78993625266729, -0.15582684282751913, 4.883087185713722},
};

// ---------------------------------------------------------------------------
// Common per-trajectory controller / loop state (all FP64)
// ---------------------------------------------------------------------------
struct Ctl {
    double t, dt, dtpropose, tprev, tf, dtmax, EEst, q11, errold;
    double tdir;   // integration direction sign(tf - t0) (+1 when tf == t0); dt and dtmax carry it
    int iter, success_iter, accept_step, next_step_tstop;
    int isout;   // the last attempt's isoutofdomain verdict (integrator.isout)
    int nf, naccept, nreject, njacs, nw, nsolve;
    int spos, send, nsaved, tf_in_saveat;
    // tstops queue (initialize_tstops, solve.jl:1019-1030): the user times
    // strictly inside (t0, tf) are P.tstops[tsp], P.tstops[tsp + dir], ... up to
    // (not including) tsend, dir = tdir (the shared list is sorted ascending),
    // then tf; tnext = the first one.  saveat likewise with spos / send.
    int tsp, tsend;
    double tnext;
    // discrete callbacks: a callback modified u after the last accepted step
    // (integrator.reeval_fsal), and the condition calls so far (stats.ncondition)
    int reeval_fsal, ncond;
    // continuous callbacks: 1 + the callback of the last event (0: none) and its
    // condition value there (integrator.event_last_time, last_event_error)
    int event_last;
    unsigned sim_mask;   // its components when a vector callback's (prev_simultaneous_events)
    double last_event_error;
    // mixed engines: fastpower(errold, beta2), kept from the accept that set
    // errold (NaN: not known yet), and log2(EEst) of the last controller call
    float erroldp, lgEEst;
};

// controller constants (alg_utils.jl:407-879; controllers.jl:744-803)
struct CtlConst {
    double qmin, qmax, qmax_first, gamma, beta1, beta2, qoldinit, qsteady_min, qsteady_max;
};
__device__ __forceinline__ CtlConst ctl_const(bool rosenbrock) {
    CtlConst c;
    c.qmin = 0.2; c.qmax = 10.0; c.qmax_first = 10000.0; c.gamma = 0.9;
    c.beta1 = 0.14; c.beta2 = 0.08; c.qoldinit = 1e-4;
    c.qsteady_min = 1.0; c.qsteady_max = rosenbrock ? 1.2 : 1.0;
    return c;
}

// integrator_utils.jl:268-324 (adaptive branch).  apply_step!
// (integrator_utils.jl:165-191) also calls it right after dt = dtpropose; of
// that call only its clip dt = min(|dt|, distance) survives (the loop header's
// fix_dt_at_bounds + modify_dt_for_tstops overwrite dtpropose and the flag), so
// the engines apply just that clip there (apply_step_tstop_clip).  It matters:
// the second call sets dtpropose from the clipped dt, and a tstop step's accept
// restores dt = dtpropose.
__device__ __forceinline__ void modify_dt_for_tstops(Ctl& c) {
    const double tstop = c.tnext;   // first queued tstop (tf last)
    // |tdir tstop - tdir t| = |tstop - t|
    double distance = fabs(tstop - c.t);
    const double tol = (isfinite(tstop) && isfinite(c.t)) ? 100.0 * jl_eps(jl_max(fabs(c.t), fabs(tstop))) : 0.0;
    double original_dt = fabs(c.dt);
    c.dtpropose = c.tdir * original_dt;
    // the tstop target of a flagged step is c.tnext (used at its accept)
    c.next_step_tstop = (original_dt + tol < distance) ? 0 : 1;
    c.dt = c.tdir * jl_min(original_dt, distance);
}
// The lasting effect of apply_step!'s modify_dt_for_tstops! (see above).
__device__ __forceinline__ void apply_step_tstop_clip(Ctl& c) {
    c.dt = c.tdir * jl_min(fabs(c.dt), fabs(c.tnext - c.t));
}
// integrator_utils.jl:1243-1256, utils.jl:90-94 (backward, dt = min(dt, |dtmin|)
// leaves a negative dt as it is, as in the reference)
__device__ __forceinline__ void fix_dt_at_bounds(Ctl& c, double dtmin) {
    const double dmin = fabs(jl_max(jl_eps(c.t), dtmin));
    if (c.tdir > 0) {
        c.dt = jl_min(c.dtmax, c.dt);
        c.dt = jl_max(c.dt, dmin);
    } else {
        c.dt = jl_max(c.dtmax, c.dt);
        c.dt = jl_min(c.dt, dmin);
    }
}
// controllers.jl:805-821.  HW: the mixed engine's form (FastPower's Float32
// log2/exp2 with the hardware reciprocal and exp2, FP64 quotients by div_nr).
template <bool HW = false>
__device__ __forceinline__ double stepsize_controller(Ctl& c, const CtlConst& k) {
    // 1/qmax as a select of two folded constants, not a division per attempt
    const double lo = (c.success_iter == 0) ? 1.0 / k.qmax_first : 1.0 / k.qmax, hi = 1.0 / k.qmin;
    double q;
    if (c.EEst == 0.0) {
        q = lo;
    } else {
        double q11, x;
        if constexpr (HW) {
            // fastpower_hw(EEst, beta1) with its log2 kept for the accept;
            // errold's power comes from the accept that set it
            c.lgEEst = fastlog2_hw(c.EEst);
            q11 = __builtin_amdgcn_exp2f((float)k.beta1 * c.lgEEst);
            if (c.erroldp != c.erroldp) c.erroldp = fastpower_hw(c.errold, k.beta2);
            x = div_nr(q11, (double)c.erroldp * k.gamma);
        } else {
            q11 = fastpower(c.EEst, k.beta1);
            q = q11 / fastpower(c.errold, k.beta2);
            x = q / k.gamma;
        }
        c.q11 = q11;
        q = (x > hi) ? hi : ((x < lo) ? lo : x);
    }
    return q;
}
// controllers.jl:823-836 + integrator_utils.jl:1199-1210
template <bool HW = false>
__device__ __forceinline__ void accept_controller(Ctl& c, const CtlConst& k, double q, double dtmin) {
    if (k.qsteady_min <= q && q <= k.qsteady_max) q = 1.0;
    c.errold = jl_max(c.EEst, k.qoldinit);
    double dtnew;
    if constexpr (HW) {
        // errold = EEst (nonzero: the controller took its log2) gives
        // fastpower_hw(errold, beta2) from that log2, bit for bit
        c.erroldp = (c.errold == c.EEst) ? __builtin_amdgcn_exp2f((float)k.beta2 * c.lgEEst)
                                         : fastpower_hw(c.errold, k.beta2);
        dtnew = div_nr(c.dt, q);   // q in [1/qmax, 1/qmin]
    } else {
        dtnew = c.dt / q;
    }
    double dp = c.tdir * jl_min(fabs(c.dtmax), fabs(dtnew));
    c.dtpropose = c.tdir * jl_max(fabs(dp), fabs(jl_max(jl_eps(c.t), dtmin)));
}
// controllers.jl:838-843 (HW: the mixed engines' Newton-refined FP64 quotients;
// a wave runs this whenever any of its lanes rejects, i.e. nearly every attempt)
template <bool HW = false>
__device__ __forceinline__ void reject_controller(Ctl& c, const CtlConst& k) {
    if constexpr (HW) c.dt = div_nr(c.dt, jl_min(1.0 / k.qmin, div_nr(c.q11, k.gamma)));
    else c.dt = c.dt / jl_min(1.0 / k.qmin, c.q11 / k.gamma);
}

template <typename S, int N>
__device__ __forceinline__ bool all_finite(const S* u) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; i++) ok = ok && isfinite((double)u[i]);
    return ok;
}

// check_error.jl:69-118 (adaptive, default dtmin/unstable_check)
__device__ __forceinline__ int check_error(const Ctl& c, const Params& P, bool u_finite) {
    if (isnan(c.dt)) return RC_DTNAN;
    if ((i64)c.iter > P.maxiters) return RC_MAXITERS;
    // t + dt < tdir first(tstops), the physical time of the first tstop, in
    // both directions (check_error.jl:91-99)
    if (fabs(c.dt) <= fabs(P.dtmin) && (!c.accept_step || c.t + c.dt < c.tnext)) return RC_DTLESSTHANMIN;
    else if (!c.accept_step && fabs(c.dt) <= fabs(jl_eps(c.t))) return RC_UNSTABLE;
    if (c.accept_step && !u_finite) return RC_UNSTABLE;
    return RC_SUCCESS;
}

// Index of the first entry of the sorted list a[0..n) that is > x (strict) or
// >= x (!strict).
__device__ __forceinline__ int sorted_bound(const double* a, int n, double x, bool strict) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (strict ? (a[m] <= x) : (a[m] < x)) lo = m + 1;
        else hi = m;
    }
    return lo;
}
// tstops cursor over the shared ascending list: the queued user tstops are the
// entries strictly inside (t, tf) in the integration direction (tsp stepping
// by tdir up to tsend), tnext the first of them, tf once they are used up.
__device__ __forceinline__ void tstops_init(Ctl& c, const Params& P, double t) {
    if (c.tdir > 0) {
        c.tsp = sorted_bound(P.tstops, P.ntstops, t, true);
        c.tsend = sorted_bound(P.tstops, P.ntstops, c.tf, false);
    } else {
        c.tsp = sorted_bound(P.tstops, P.ntstops, t, false) - 1;
        c.tsend = sorted_bound(P.tstops, P.ntstops, c.tf, true) - 1;
    }
    if (c.tdir > 0 ? (c.tsend < c.tsp) : (c.tsend > c.tsp)) c.tsend = c.tsp;
    c.tnext = (c.tsp != c.tsend) ? P.tstops[c.tsp] : c.tf;
}
// handle_tstop! (integrator_utils.jl:1290-1314) after an accepted step: once t
// reaches the first tstop every queued copy of it is popped (tf ends the loop).
__device__ __forceinline__ void handle_tstop(Ctl& c, const Params& P) {
    if (c.tsp != c.tsend && c.t == c.tnext) {
        const int dir = (c.tdir > 0) ? 1 : -1;
        while (c.tsp != c.tsend && P.tstops[c.tsp] == c.t) c.tsp += dir;
        c.tnext = (c.tsp != c.tsend) ? P.tstops[c.tsp] : c.tf;
    }
}

// Saveat cursor setup (initialize_saveat, solve.jl:1103-1123): the entries of
// the shared ascending list with tdir t0 < tdir s <= tdir tf, in the
// integration direction (spos stepping by tdir up to send).
__device__ __forceinline__ void saveat_init(Ctl& c, const Params& P, double t0) {
    if (c.tdir > 0) {
        c.spos = sorted_bound(P.saveat, P.nsaveat, t0, true);
        c.send = sorted_bound(P.saveat, P.nsaveat, c.tf, true);
    } else {
        c.spos = sorted_bound(P.saveat, P.nsaveat, t0, false) - 1;
        c.send = sorted_bound(P.saveat, P.nsaveat, c.tf, false) - 1;
    }
    if (c.tdir > 0 ? (c.send < c.spos) : (c.send > c.spos)) c.send = c.spos;
    const int k = sorted_bound(P.saveat, P.nsaveat, c.tf, false);
    c.tf_in_saveat = (k < P.nsaveat && P.saveat[k] == c.tf) ? 1 : 0;
}

template <int N, typename V>
__device__ __forceinline__ void save_point(Ctl& c, const Params& P, i64 j, double t, const V* v) {
    int s = c.nsaved;
    if (s < P.max_saves) {
        P.t_out[(i64)s * P.N + j] = t;
#pragma unroll
        for (int i = 0; i < N; i++) P.u_out[((i64)s * P.nmax + i) * P.N + j] = (double)v[i];
    }
    c.nsaved = s + 1;
}

// handle_callbacks! (integrator_utils.jl:1081-1132) for a CallbackSet of
// DiscreteCallbacks, after an accepted step's dt proposal.  Each callback is
// applied as apply_discrete_callback! does (lib/DiffEqBase/src/callbacks.jl:
// 645-683): when its condition holds at (u, t), savevalues! (saveat points,
// save_everystep), a forced save of the pre-affect state unless savevalues!
// saved at exactly t, affect!, then a forced save of the new state; the
// modified state makes the next accept re-evaluate the FSAL stage
// (reeval_internals_due_to_modification!).  Every condition call counts in
// stats.ncondition.  `savevalues` is the method's savevalues! and returns
// savedexactly.  The result is true when an affect! called terminate!.
#ifdef ODEMI_NDCB
template <int K, int N, typename S, typename Q, typename SaveFn>
__device__ __forceinline__ void dcb_apply(Ctl& c, const Params& P, i64 j, S* u, const Q* p, SaveFn& savevalues,
                                          bool& saved_in_cb, bool& term) {
    if constexpr (K < ODEMI_NDCB) {
        c.ncond++;
        if (OdemiDcb<K>::template cond<S, Q>(u, p, (S)c.t)) {
            const bool savedexactly = savevalues();
            saved_in_cb = true;
            if (OdemiDcb<K>::save0 && !savedexactly) save_point<N>(c, P, j, c.t, u);
            OdemiDcb<K>::template affect<S, Q>(u, p, (S)c.t);
            c.reeval_fsal = 1;
            if (OdemiDcb<K>::term) term = true;   // terminate!: the tstops queue is emptied
            if (OdemiDcb<K>::save1) save_point<N>(c, P, j, c.t, u);
        }
        dcb_apply<K + 1, N>(c, P, j, u, p, savevalues, saved_in_cb, term);
    }
}
#endif
// `continuous(saved_in_cb)` runs the method's continuous callbacks first
// (find_first_continuous_callback + apply_callback!) and returns terminate!.
template <int N, typename S, typename Q, typename SaveFn, typename ContFn>
__device__ __forceinline__ bool handle_callbacks(Ctl& c, const Params& P, i64 j, S* u, const Q* p,
                                                 SaveFn&& savevalues, ContFn&& continuous) {
#if defined(ODEMI_NDCB) || defined(ODEMI_NCCB)
    bool saved_in_cb = false, term = continuous(saved_in_cb);
#ifdef ODEMI_NDCB
    dcb_apply<0, N>(c, P, j, u, p, savevalues, saved_in_cb, term);
#endif
    if (!saved_in_cb) savevalues();
    return term;
#else
    (void)c; (void)P; (void)j; (void)u; (void)p; (void)continuous;
    savevalues();
    return false;
#endif
}

// RMS norm of residuals, FP64 accumulation (calculate_residuals.jl:9-14 +
// common_defaults.jl:69-75)
template <int N, typename E, typename S>
__device__ __forceinline__ double residual_norm(const E* ut, const S* u0, const S* u1,
                                                double abstol, double reltol) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double den = fma(max_fast(fabs((double)u0[i]), fabs((double)u1[i])), reltol, abstol);
        double r = (double)ut[i] / den;
        s = fma(r, r, s);
    }
    return sqrt(s / (double)(N > 1 ? N : 1));
}

// The mixed engine's FP64 error norm: the same sum as residual_norm with the
// Newton-refined FP64 quotient and square root.
template <int N>
__device__ __forceinline__ double residual_norm_nr(const double* ut, const double* u0, const double* u1,
                                                   double abstol, double reltol) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double den = fma(max_fast(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
        double r;
        if (__builtin_expect(!(den >= 0x1p-1000), 0)) r = ut[i] / den;   // abstol ~ 0 and u ~ 0
        else r = div_nr(ut[i], den);
        s = fma(r, r, s);
    }
    return sqrt_nr(s * (1.0 / (double)(N > 1 ? N : 1)));
}

template <int N>
__device__ __forceinline__ double ode_norm(const double* x) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) s = fma(x[i], x[i], s);
    return sqrt(s / (double)(N > 1 ? N : 1));
}

// initdt.jl:17-298 restated for a lane-resident state (FP64 evaluation).
template <int N>
__device__ double initdt_lane(const Ctl& c, const double* u0, const double* p, double reltol,
                              double abstol, double dtmax_opt, double dtmin_opt) {
    const double t = c.t, tdir = c.tdir;
    const double dtmax_tdir = jl_min(fabs(dtmax_opt), fabs(c.tnext - t));   // first tstop
    const double dtmin = nextfloat_pos(jl_max(dtmin_opt, jl_eps(t)));
    const double smalldt = jl_max(dtmin, 1e-6);
    if (N == 0) return tdir * jl_max(smalldt, dtmin);
    double sk[N], f0[N], f1[N], tmp[N];
#pragma unroll
    for (int i = 0; i < N; i++) sk[i] = fma(fabs(u0[i]), reltol, abstol);
    odemi_f<double>(f0, u0, p, t);
#pragma unroll
    for (int i = 0; i < N; i++) tmp[i] = u0[i] / sk[i];
    double d0 = ode_norm<N>(tmp);
#pragma unroll
    for (int i = 0; i < N; i++) tmp[i] = f0[i] / sk[i] * 1.0;
    double d1 = ode_norm<N>(tmp);
    if (isnan(d1)) return tdir * dtmin;
    double dt0 = ((d0 < 1e-5) | (d1 < 1e-5)) ? smalldt : (d0 / d1) / 100.0;
    dt0 = jl_min(dt0, dtmax_tdir);
    if (dt0 < 10.0 * 2.220446049250313e-16) return tdir * jl_max(smalldt, dtmin);
    const double dt0_tdir = tdir * dt0;
#pragma unroll
    for (int i = 0; i < N; i++) tmp[i] = fma(dt0_tdir, f0[i], u0[i]);
    odemi_f<double>(f1, tmp, p, t + dt0_tdir);
    bool same = true;
#pragma unroll
    for (int i = 0; i < N; i++) same = same && (f0[i] == f1[i]);
    if (same) return tdir * jl_max(dtmin, 100.0 * dt0);
#pragma unroll
    for (int i = 0; i < N; i++) tmp[i] = (f1[i] - f0[i]) / sk[i] * 1.0;
    double d2 = ode_norm<N>(tmp) / dt0 * 1.0;
    double mx = jl_max(d1, d2);
    double dt1;
    if (mx < 1e-15) dt1 = jl_max(1e-6, dt0 * 0.001);
    else dt1 = det_exp(-(2.0 + det_log(mx) / kLN10) / 5.0 * kLN10);
    return tdir * jl_max(dtmin, jl_min(jl_min(100.0 * dt0, dt1), dtmax_tdir));
}

// handle_dt! (solve.jl:968-985): a positive user dt of a backward solve takes
// tdir's sign.
__device__ __forceinline__ double user_dt(const Ctl& c, const Params& P) {
    return (P.dt_user > 0.0 && c.tdir < 0) ? c.tdir * P.dt_user : P.dt_user;
}

// Common per-trajectory initialisation of the controller state.
__device__ __forceinline__ void ctl_init_t(Ctl& c, const Params& P, double t0, double tf) {
    c.t = t0; c.tprev = t0; c.tf = tf;
    c.tdir = (tf < t0) ? -1.0 : 1.0;
    // dtmax defaults to tf - t0; a positive one takes tdir's sign (solve.jl:401)
    c.dtmax = (P.dtmax > 0.0) ? c.tdir * P.dtmax : (tf - t0);
    c.EEst = 1.0; c.q11 = 1.0; c.errold = 1e-4;
    c.iter = 0; c.success_iter = 0; c.accept_step = 0; c.next_step_tstop = 0; c.isout = 0;
    c.reeval_fsal = 0; c.ncond = 0; c.event_last = 0; c.sim_mask = 0u; c.last_event_error = 0.0;
    c.erroldp = __builtin_nanf(""); c.lgEEst = 0.0f;
    c.dt = 0.0; c.dtpropose = 0.0;
    c.nf = 0; c.naccept = 0; c.nreject = 0; c.njacs = 0; c.nw = 0; c.nsolve = 0;
    c.nsaved = 0; c.spos = 0; c.send = 0; c.tf_in_saveat = 0;
    if (P.nsaveat > 0) saveat_init(c, P, t0);
    c.tsp = 0; c.tsend = 0; c.tnext = tf;
    if (P.ntstops > 0) tstops_init(c, P, t0);
}
__device__ __forceinline__ void ctl_init(Ctl& c, const Params& P, i64 j) { ctl_init_t(c, P, P.t0[j], P.tf[j]); }

// solution_endpoint_match_cur_integrator! (integrator_utils.jl:540-588) and
// result write-out.
template <int N, typename S>
__device__ void finish_lane(Ctl& c, const Params& P, i64 j, const S* u, int rc) {
    if ((rc == RC_SUCCESS || rc == RC_TERMINATED) && P.save_end) {   // postamble! runs after terminate! too
        bool doit = (c.nsaved == 0);
        if (!doit) {
            int s = c.nsaved - 1;
            double last = (s < P.max_saves) ? P.t_out[(i64)s * P.N + j] : __longlong_as_double(0x7ff8000000000000LL);
            doit = (last != c.t) && (P.save_end_user == 1 || c.tf_in_saveat || c.t == c.tf || P.nsaveat == 0);
        }
        if (doit) save_point<N>(c, P, j, c.t, u);
    }
    if (P.u_final) {
#pragma unroll
        for (int i = 0; i < N; i++) P.u_final[(i64)i * P.N + j] = (double)u[i];
    }
    if (P.t_final) P.t_final[j] = c.t;
    if (P.nsaved) P.nsaved[j] = c.nsaved;
    P.retcode[j] = rc;
    if (P.stats) {
        P.stats[ST_NF * P.N + j] = c.nf;
        P.stats[ST_NACCEPT * P.N + j] = c.naccept;
        P.stats[ST_NREJECT * P.N + j] = c.nreject;
        P.stats[ST_NJACS * P.N + j] = c.njacs;
        P.stats[ST_NW * P.N + j] = c.nw;
        P.stats[ST_NSOLVE * P.N + j] = c.nsolve;
        P.stats[ST_NCONDITION * P.N + j] = c.ncond;
    }
}

// Queue-order records of the final-state-only Tsit5 persistent engine.  The
// lanes take trajectories in longest-first queue order, so trajectory j's
// inputs and results sit at scattered indices of the SoA arrays: read and
// written in place they cost a 64-128 B line per 8-byte value (1.7 GB per
// launch for a 1 M Lorenz batch against 0.2 GB of algorithmic bytes).  With
// records the prep kernel (one thread per trajectory, coalesced SoA reads)
// writes trajectory j's inputs and start-up results as one record at its
// queue slot inv[j]; a lane reads the record of the slot it takes (a refill
// batch takes consecutive slots: contiguous records) and writes its result
// record whole at that slot; rec_unpack (one thread per trajectory) moves the
// results to the SoA outputs with coalesced writes.  Input record, doubles:
// u0[N] p[NP] t0 tf k1[N] dt.  Result record: u_final[N] t_final, then int32
// retcode stats[7] (64 B for N = 3: one aligned sector per record).  Used only
// without saves before the end point (no save_start, saveat, save_everystep
// or callbacks): the end point's saved copy is the final state, and nsaved is
// 1 exactly when finish_lane saves it.
constexpr int rec_in_words(int n, int np) { return (2 * n + np + 3 + 1) & ~1; }
constexpr int rec_out_words(int n) { return (n + 1 + 4 + 1) & ~1; }

template <int N, typename S>
__device__ __forceinline__ void finish_lane_rec(const Ctl& c, const Params& P, i64 slot, const S* u, int rc) {
    double* r = P.rec_out + slot * rec_out_words(N);
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = (double)u[i];
    r[N] = c.t;
    int* q = (int*)(r + N + 1);
    q[0] = rc;
    q[1 + ST_NF] = c.nf; q[1 + ST_NACCEPT] = c.naccept; q[1 + ST_NREJECT] = c.nreject;
    q[1 + ST_NJACS] = c.njacs; q[1 + ST_NW] = c.nw; q[1 + ST_NSOLVE] = c.nsolve;
    q[1 + ST_NCONDITION] = c.ncond;
}
template <int N>
__device__ void rec_unpack(const Params& P) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    const double* r = P.rec_out + (i64)P.inv[j] * rec_out_words(N);
    const int* q = (const int*)(r + N + 1);
    const int rc = q[0];
    const int ns = ((rc == RC_SUCCESS || rc == RC_TERMINATED) && P.save_end) ? 1 : 0;   // finish_lane's end point
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (P.u_final) P.u_final[(i64)i * P.N + j] = r[i];
        if (ns > 0 && P.max_saves > 0) P.u_out[(i64)i * P.N + j] = r[i];   // save row 0
    }
    if (P.t_final) P.t_final[j] = r[N];
    if (ns > 0 && P.max_saves > 0) P.t_out[j] = r[N];
    if (P.nsaved) P.nsaved[j] = ns;
    P.retcode[j] = rc;
    if (P.stats) {
#pragma unroll
        for (int k = 0; k < ST_COUNT; k++) P.stats[(i64)k * P.N + j] = q[1 + k];
    }
}

// ---------------------------------------------------------------------------
// Tsit5, one trajectory per lane.
//   S = state type (uprev/u accumulation), K = stage type (RHS, k_i, stage sums)
//   FP64 engine: S = K = double.  Mixed engine: S = double, K = float.
// ---------------------------------------------------------------------------
template <typename S, typename K, int N>
struct Tsit5Lane {
    Ctl c;
    S u[N], uprev[N];
    K k1[N], k2[N], k3[N], k4[N], k5[N], k6[N], k7[N];
    K p[ODEMI_NP];
};

// Loads u0 and p of trajectory j into the lane (FP64 copies in u0d / pd).
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_load(Tsit5Lane<S, K, N>& L, const Params& P, i64 j, double (&u0d)[N],
                                           double (&pd)[ODEMI_NP]) {
    ctl_init(L.c, P, j);
#pragma unroll
    for (int i = 0; i < N; i++) { u0d[i] = P.u0[(i64)i * P.N + j]; L.u[i] = (S)u0d[i]; L.uprev[i] = L.u[i]; }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) { pd[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0; L.p[k] = (K)pd[k]; }
}
// initialize!: FSAL pre-start k1 = f(u0), then the initial dt.
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_first_stage_dt(Tsit5Lane<S, K, N>& L, const Params& P, const double (&u0d)[N],
                                                     const double (&pd)[ODEMI_NP]) {
    K uk[N];
#pragma unroll
    for (int i = 0; i < N; i++) uk[i] = (K)L.uprev[i];
    odemi_f<K>(L.k1, uk, L.p, (K)L.c.t);
    L.c.nf += 1;
    if (P.dt_user != 0.0) {  // user dt (a NaN dt is reported as DtNaN)
        L.c.dt = user_dt(L.c, P);
    } else {
        L.c.dt = initdt_lane<N>(L.c, u0d, pd, P.reltol, P.abstol, L.c.dtmax, P.dtmin);
        L.c.nf += 2;
    }
}
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_init(Tsit5Lane<S, K, N>& L, const Params& P, i64 j, i64 slot = -1) {
    if (P.rec_in) {   // the record prep wrote for this queue slot
        const double* r = P.rec_in + slot * rec_in_words(N, ODEMI_NP);
        ctl_init_t(L.c, P, r[N + ODEMI_NP], r[N + ODEMI_NP + 1]);
#pragma unroll
        for (int i = 0; i < N; i++) { L.u[i] = (S)r[i]; L.uprev[i] = L.u[i]; }
#pragma unroll
        for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (K)r[N + k];
#pragma unroll
        for (int i = 0; i < N; i++) L.k1[i] = (K)r[N + ODEMI_NP + 2 + i];
        L.c.dt = r[2 * N + ODEMI_NP + 2];
        L.c.nf += (P.dt_user != 0.0) ? 1 : 3;
        L.c.dtpropose = L.c.dt;
        return;
    }
    double u0d[N], pd[ODEMI_NP];
    tsit5_load(L, P, j, u0d, pd);
    if (P.save_start) save_point<N>(L.c, P, j, L.c.t, L.u);
    if (P.pre_dt) {
        // computed by tsit5_prep with the same arithmetic
#pragma unroll
        for (int i = 0; i < N; i++) L.k1[i] = (K)P.pre_k1[(i64)i * P.N + j];
        L.c.dt = P.pre_dt[j];
        L.c.nf += (P.dt_user != 0.0) ? 1 : 3;
    } else {
        tsit5_first_stage_dt(L, P, u0d, pd);
    }
    L.c.dtpropose = L.c.dt;
}
// One thread per trajectory, ahead of the persistent kernel: the start-up work
// (one RHS in the stage precision, initdt's two FP64 RHS and norms) runs fully
// parallel instead of stalling the other lanes of a wave whenever one lane
// picks up a new trajectory.
template <typename S, typename K, int N>
__device__ void tsit5_prep(const Params& P) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    Tsit5Lane<S, K, N> L;
    double u0d[N], pd[ODEMI_NP];
    tsit5_load(L, P, j, u0d, pd);
    tsit5_first_stage_dt(L, P, u0d, pd);
    if (P.rec_in) {
        double* r = P.rec_in + (i64)P.inv[j] * rec_in_words(N, ODEMI_NP);
#pragma unroll
        for (int i = 0; i < N; i++) r[i] = u0d[i];
#pragma unroll
        for (int k = 0; k < ODEMI_NP; k++) r[N + k] = pd[k];
        r[N + ODEMI_NP] = L.c.t;
        r[N + ODEMI_NP + 1] = L.c.tf;
#pragma unroll
        for (int i = 0; i < N; i++) r[N + ODEMI_NP + 2 + i] = (double)L.k1[i];
        r[2 * N + ODEMI_NP + 2] = L.c.dt;
        return;
    }
#pragma unroll
    for (int i = 0; i < N; i++) P.pre_k1[(i64)i * P.N + j] = (double)L.k1[i];
    P.pre_dt[j] = L.c.dt;
}

// Mixed engine (K = float): stage inputs are FP32 increments on an FP32 copy of
// uprev and the embedded error combination sum(btilde_i k_i) is FP32 like the
// other stage sums; the solution update u = uprev + dt * sum(a7j kj), t + dt,
// the error norm and the controller stay FP64 (Newton-refined hardware
// reciprocal and square root instead of the IEEE sequences).  On CDNA4 FP32
// and FP64 VALU ops issue at one rate, so the gain is in the operations dropped
// (per-stage conversions, FastPower's exp2 through the exact FP64 det_exp, the
// IEEE division and square-root fix-up sequences).
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_perform_step_mixed(Tsit5Lane<S, K, N>& L, const Params& P) {
    typedef Tsit5Tab T;
    const double t = L.c.t, dt = L.c.dt;
    const float dtf = (float)dt;
    float ub[N], e[N], tmp[N];
    const float a = (float)(dt * T::a21);
#pragma unroll
    for (int i = 0; i < N; i++) {
        ub[i] = (float)L.uprev[i];
        tmp[i] = fmaf(a, L.k1[i], ub[i]);
        e[i] = (float)T::bt1 * L.k1[i];
    }
    odemi_f<float>(L.k2, tmp, L.p, (float)fma(T::c1, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        float acc = (float)T::a31 * L.k1[i];
        acc = fmaf((float)T::a32, L.k2[i], acc);
        tmp[i] = fmaf(dtf, acc, ub[i]);
        e[i] = fmaf((float)T::bt2, L.k2[i], e[i]);
    }
    odemi_f<float>(L.k3, tmp, L.p, (float)fma(T::c2, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        float acc = (float)T::a41 * L.k1[i];
        acc = fmaf((float)T::a42, L.k2[i], acc);
        acc = fmaf((float)T::a43, L.k3[i], acc);
        tmp[i] = fmaf(dtf, acc, ub[i]);
        e[i] = fmaf((float)T::bt3, L.k3[i], e[i]);
    }
    odemi_f<float>(L.k4, tmp, L.p, (float)fma(T::c3, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        float acc = (float)T::a51 * L.k1[i];
        acc = fmaf((float)T::a52, L.k2[i], acc);
        acc = fmaf((float)T::a53, L.k3[i], acc);
        acc = fmaf((float)T::a54, L.k4[i], acc);
        tmp[i] = fmaf(dtf, acc, ub[i]);
        e[i] = fmaf((float)T::bt4, L.k4[i], e[i]);
    }
    odemi_f<float>(L.k5, tmp, L.p, (float)fma(T::c4, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        float acc = (float)T::a61 * L.k1[i];
        acc = fmaf((float)T::a62, L.k2[i], acc);
        acc = fmaf((float)T::a63, L.k3[i], acc);
        acc = fmaf((float)T::a64, L.k4[i], acc);
        acc = fmaf((float)T::a65, L.k5[i], acc);
        tmp[i] = fmaf(dtf, acc, ub[i]);
        e[i] = fmaf((float)T::bt5, L.k5[i], e[i]);
    }
    odemi_f<float>(L.k6, tmp, L.p, (float)(t + dt));
#pragma unroll
    for (int i = 0; i < N; i++) {
        float acc = (float)T::a71 * L.k1[i];
        acc = fmaf((float)T::a72, L.k2[i], acc);
        acc = fmaf((float)T::a73, L.k3[i], acc);
        acc = fmaf((float)T::a74, L.k4[i], acc);
        acc = fmaf((float)T::a75, L.k5[i], acc);
        acc = fmaf((float)T::a76, L.k6[i], acc);
        L.u[i] = fma(dt, (double)acc, L.uprev[i]);
        e[i] = fmaf((float)T::bt6, L.k6[i], e[i]);
        tmp[i] = (float)L.u[i];
    }
    odemi_f<float>(L.k7, tmp, L.p, (float)(t + dt));
    L.c.nf += 6;
    double ea[N];   // utilde = dt * e, then the FP64 norm
#pragma unroll
    for (int i = 0; i < N; i++) ea[i] = dt * (double)fmaf((float)T::bt7, L.k7[i], e[i]);
    L.c.EEst = residual_norm_nr<N>(ea, L.uprev, L.u, P.abstol, P.reltol);
}

template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_perform_step(Tsit5Lane<S, K, N>& L, const Params& P) {
    if constexpr (sizeof(K) < sizeof(S)) {
        tsit5_perform_step_mixed(L, P);
        return;
    }
    typedef Tsit5Tab T;
    const double t = L.c.t, dt = L.c.dt;
    K tmp[N];
    double ea[N];  // embedded error sum bt1*k1 + ... + bt7*k7, accumulated in FP64 as stages appear
    const K a = (K)(dt * T::a21);
#pragma unroll
    for (int i = 0; i < N; i++) {
        tmp[i] = (K)fma((S)a, (S)L.k1[i], L.uprev[i]);
        ea[i] = T::bt1 * (double)L.k1[i];
    }
    odemi_f<K>(L.k2, tmp, L.p, (K)fma(T::c1, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a31 * L.k1[i];
        acc = fma((K)T::a32, L.k2[i], acc);
        tmp[i] = (K)fma((S)dt, (S)acc, L.uprev[i]);
        ea[i] = fma(T::bt2, (double)L.k2[i], ea[i]);
    }
    odemi_f<K>(L.k3, tmp, L.p, (K)fma(T::c2, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a41 * L.k1[i];
        acc = fma((K)T::a42, L.k2[i], acc);
        acc = fma((K)T::a43, L.k3[i], acc);
        tmp[i] = (K)fma((S)dt, (S)acc, L.uprev[i]);
        ea[i] = fma(T::bt3, (double)L.k3[i], ea[i]);
    }
    odemi_f<K>(L.k4, tmp, L.p, (K)fma(T::c3, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a51 * L.k1[i];
        acc = fma((K)T::a52, L.k2[i], acc);
        acc = fma((K)T::a53, L.k3[i], acc);
        acc = fma((K)T::a54, L.k4[i], acc);
        tmp[i] = (K)fma((S)dt, (S)acc, L.uprev[i]);
        ea[i] = fma(T::bt4, (double)L.k4[i], ea[i]);
    }
    odemi_f<K>(L.k5, tmp, L.p, (K)fma(T::c4, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a61 * L.k1[i];
        acc = fma((K)T::a62, L.k2[i], acc);
        acc = fma((K)T::a63, L.k3[i], acc);
        acc = fma((K)T::a64, L.k4[i], acc);
        acc = fma((K)T::a65, L.k5[i], acc);
        tmp[i] = (K)fma((S)dt, (S)acc, L.uprev[i]);
        ea[i] = fma(T::bt5, (double)L.k5[i], ea[i]);
    }
    odemi_f<K>(L.k6, tmp, L.p, (K)(t + dt));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a71 * L.k1[i];
        acc = fma((K)T::a72, L.k2[i], acc);
        acc = fma((K)T::a73, L.k3[i], acc);
        acc = fma((K)T::a74, L.k4[i], acc);
        acc = fma((K)T::a75, L.k5[i], acc);
        acc = fma((K)T::a76, L.k6[i], acc);
        L.u[i] = fma((S)dt, (S)acc, L.uprev[i]);
        ea[i] = fma(T::bt6, (double)L.k6[i], ea[i]);
    }
#pragma unroll
    for (int i = 0; i < N; i++) tmp[i] = (K)L.u[i];
    odemi_f<K>(L.k7, tmp, L.p, (K)(t + dt));
    L.c.nf += 6;
#pragma unroll
    for (int i = 0; i < N; i++) ea[i] = dt * fma(T::bt7, (double)L.k7[i], ea[i]);
    L.c.EEst = residual_norm<N>(ea, L.uprev, L.u, P.abstol, P.reltol);
}

// interpolants.jl:52-64 (@evalpoly Horner via muladd), FP64 evaluation.
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_interp(const Tsit5Lane<S, K, N>& L, double th, double dt, double* out) {
    double th2 = th * th;
    double b[7];
    b[0] = th * fma(th, fma(th, fma(th, kTsitR[0][3], kTsitR[0][2]), kTsitR[0][1]), kTsitR[0][0]);
#pragma unroll
    for (int s = 1; s < 7; s++) b[s] = th2 * fma(th, fma(th, kTsitR[s][3], kTsitR[s][2]), kTsitR[s][1]);
#pragma unroll
    for (int i = 0; i < N; i++) {
        double acc = (double)L.k1[i] * b[0];
        acc = fma((double)L.k2[i], b[1], acc);
        acc = fma((double)L.k3[i], b[2], acc);
        acc = fma((double)L.k4[i], b[3], acc);
        acc = fma((double)L.k5[i], b[4], acc);
        acc = fma((double)L.k6[i], b[5], acc);
        acc = fma((double)L.k7[i], b[6], acc);
        out[i] = fma(dt, acc, (double)L.uprev[i]);
    }
}

// savevalues! (integrator_utils.jl:340-414); returns savedexactly.
template <typename S, typename K, int N>
__device__ __forceinline__ bool tsit5_savevalues(Tsit5Lane<S, K, N>& L, const Params& P, i64 j) {
    Ctl& c = L.c;
    bool exact = false;
    while (c.spos != c.send && c.tdir * P.saveat[c.spos] <= c.tdir * c.t) {
        double curt = P.saveat[c.spos];
        c.spos += (c.tdir > 0) ? 1 : -1;
        if (curt != c.t) {
            double th = (curt - c.tprev) / c.dt;
            double v[N];
            tsit5_interp(L, th, c.dt, v);
            save_point<N>(c, P, j, curt, v);
        } else {
            if (curt == c.tf && !P.save_end) continue;
            save_point<N>(c, P, j, c.t, L.u);
            exact = true;
        }
    }
    if (P.save_everystep) {
        int s = c.nsaved;
        double last = (s > 0 && s - 1 < P.max_saves) ? P.t_out[(i64)(s - 1) * P.N + j] : __longlong_as_double(0x7ff8000000000000LL);
        if (s == 0 || ((c.t != last || c.dt == 0.0) && (P.save_end || c.t != c.tf))) {
            save_point<N>(c, P, j, c.t, L.u);
            exact = true;
        }
    }
    return exact;
}

#ifdef ODEMI_NCCB
// ---------------------------------------------------------------------------
// ContinuousCallbacks on the lanes (lib/DiffEqBase/src/callbacks.jl:157-640;
// OdemiCcb<K> holds the traced condition / affect! / affect_neg! and the
// options as constants).  Generic over the lane type through lane_interp and
// lane_addsteps.
// ---------------------------------------------------------------------------
// Julia's sign(): -1, 0, 1, NaN for NaN
__device__ __forceinline__ double jl_sign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : (x == 0 ? 0.0 : x)); }

// _ode_addsteps! for Tsit5Cache (tsit_perform_step.jl:1-38, @muladd) with
// always_calc_begin: stages k2..k7 of the step from uprev at tprev with the
// current dt, k1 kept (reeval_internals_due_to_modification! after
// change_t_via_interpolation!).  No stats change.  The mixed engine uses its
// FP32 stage forms.
template <typename S, typename K, int N>
__device__ void tsit5_addsteps(Tsit5Lane<S, K, N>& L) {
    typedef Tsit5Tab T;
    const double t = L.c.tprev, dt = L.c.dt;
    constexpr bool MIX = sizeof(K) < sizeof(S);
    K tmp[N];
    float ub[N];
    const float dtf = (float)dt;
#pragma unroll
    for (int i = 0; i < N; i++) {
        ub[i] = (float)L.uprev[i];
        if constexpr (MIX) tmp[i] = fmaf(dtf, (float)T::a21 * L.k1[i], ub[i]);
        else tmp[i] = fma(dt, T::a21 * L.k1[i], L.uprev[i]);
    }
    odemi_f<K>(L.k2, tmp, L.p, (K)fma(T::c1, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a31 * L.k1[i];
        acc = fma((K)T::a32, L.k2[i], acc);
        if constexpr (MIX) tmp[i] = fmaf(dtf, acc, ub[i]);
        else tmp[i] = fma(dt, acc, L.uprev[i]);
    }
    odemi_f<K>(L.k3, tmp, L.p, (K)fma(T::c2, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a41 * L.k1[i];
        acc = fma((K)T::a42, L.k2[i], acc);
        acc = fma((K)T::a43, L.k3[i], acc);
        if constexpr (MIX) tmp[i] = fmaf(dtf, acc, ub[i]);
        else tmp[i] = fma(dt, acc, L.uprev[i]);
    }
    odemi_f<K>(L.k4, tmp, L.p, (K)fma(T::c3, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a51 * L.k1[i];
        acc = fma((K)T::a52, L.k2[i], acc);
        acc = fma((K)T::a53, L.k3[i], acc);
        acc = fma((K)T::a54, L.k4[i], acc);
        if constexpr (MIX) tmp[i] = fmaf(dtf, acc, ub[i]);
        else tmp[i] = fma(dt, acc, L.uprev[i]);
    }
    odemi_f<K>(L.k5, tmp, L.p, (K)fma(T::c4, dt, t));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a61 * L.k1[i];
        acc = fma((K)T::a62, L.k2[i], acc);
        acc = fma((K)T::a63, L.k3[i], acc);
        acc = fma((K)T::a64, L.k4[i], acc);
        acc = fma((K)T::a65, L.k5[i], acc);
        if constexpr (MIX) tmp[i] = fmaf(dtf, acc, ub[i]);
        else tmp[i] = fma(dt, acc, L.uprev[i]);
    }
    odemi_f<K>(L.k6, tmp, L.p, (K)(t + dt));
#pragma unroll
    for (int i = 0; i < N; i++) {
        K acc = (K)T::a71 * L.k1[i];
        acc = fma((K)T::a72, L.k2[i], acc);
        acc = fma((K)T::a73, L.k3[i], acc);
        acc = fma((K)T::a74, L.k4[i], acc);
        acc = fma((K)T::a75, L.k5[i], acc);
        acc = fma((K)T::a76, L.k6[i], acc);
        tmp[i] = (K)fma(dt, (double)acc, (double)L.uprev[i]);
    }
    odemi_f<K>(L.k7, tmp, L.p, (K)(t + dt));
}

// The method's dense output at theta of the current step and its event
// re-step; the Rodas5P overloads follow its lane.
template <typename S, typename K, int N>
__device__ __forceinline__ void lane_interp(const Tsit5Lane<S, K, N>& L, double th, double* v) {
    tsit5_interp(L, th, L.c.dt, v);
}
template <typename S, typename K, int N>
__device__ __forceinline__ void lane_addsteps(Tsit5Lane<S, K, N>& L, const Params& P) {
    (void)P;
    tsit5_addsteps(L);
}

// get_condition (callbacks.jl:91-146): u at t, uprev at tprev, else the
// step's dense output; every call counts in stats.ncondition.
template <int CB, int N, typename Lane>
__device__ __forceinline__ double ccb_condition(Lane& L, double abst) {
    Ctl& c = L.c;
    c.ncond++;
    if (abst == c.t) return (double)OdemiCcb<CB>::template cond<double>(L.u, L.p, abst);
    if (abst == c.tprev) return (double)OdemiCcb<CB>::template cond<double>(L.uprev, L.p, abst);
    double v[N];
    lane_interp(L, (abst - c.tprev) / c.dt, v);
    return (double)OdemiCcb<CB>::template cond<double>(v, L.p, abst);
}

// is_event_occurrence (callbacks.jl:525-530)
template <int CB>
__device__ __forceinline__ bool ccb_is_event(double prev_sign, double next_sign) {
    return ((prev_sign < 0 && OdemiCcb<CB>::has_affect) || (prev_sign > 0 && OdemiCcb<CB>::has_affect_neg)) &&
           prev_sign * next_sign <= 0;
}

// find_callback_time (callbacks.jl:364-412) with nudge_tprev (414-426),
// check_event_occurrence (428-451; the interp_points checks at
// tprev + (i - 1) h, h = (t - tprev) / (interp_points - 1), t last) and
// find_root (480-497): the sign change bracketed down to adjacent floats by
// bisection (the interval solve with abstol = reltol = 0), the bottom-side
// float for LeftRootFind, the other for RightRootFind.  Bisection in FP64
// halves the bracket each call, so the loop ends within ~2100 calls for any
// finite bracket.
template <int CB, int N, typename Lane>
__device__ bool ccb_find_time(Lane& L, double& cb_t, double& bottom_sign, double& residual) {
    typedef OdemiCcb<CB> C;
    Ctl& c = L.c;
    double bottom_t = c.tprev;
    double bc = ccb_condition<CB, N>(L, bottom_t);
    if (c.event_last == CB + 1) {
        if (fabs(bc - c.last_event_error) <= C::abstol) bottom_t = c.tprev + c.dt * C::repeat_nudge;
        bc = ccb_condition<CB, N>(L, bottom_t);
    }
    const double bs = jl_sign(bc);
    double top_t = c.t;
    double ts = jl_sign(ccb_condition<CB, N>(L, top_t));
    bool ev = ccb_is_event<CB>(bs, ts);
    if (C::interp_points != 0 && !ev) {
        const double h = (c.t - c.tprev) / (double)(C::interp_points - 1);
        for (int i = 2; i <= C::interp_points; i++) {
            top_t = (i == C::interp_points) ? c.t : fma((double)(i - 1), h, c.tprev);
            ts = jl_sign(ccb_condition<CB, N>(L, top_t));
            ev = ccb_is_event<CB>(bs, ts);
            if (ev) break;
        }
    }
    bottom_sign = bs;
    if (!ev) { cb_t = c.t; residual = 0.0; return false; }
    if (C::rootfind == 2 || ts == 0.0) { cb_t = top_t; residual = 0.0; return true; }
    double l = bottom_t, r = top_t;
    for (int it = 0; it < 2200; it++) {
        const double m = l + 0.5 * (r - l);
        if (m == l || m == r) break;
        const double sm = jl_sign(ccb_condition<CB, N>(L, m));
        if (sm == 0.0) { l = m; r = m; break; }
        if (sm == bs) l = m;
        else r = m;
    }
    cb_t = (C::rootfind == 1) ? r : l;
    residual = ccb_condition<CB, N>(L, cb_t);
    return true;
}

// apply_callback! (callbacks.jl:557-640): set_proposed_dt! (dtrelax = 1),
// change_t_via_interpolation! (integrator_interface.jl:5-40: u from the dense
// output, t, dt = t - tprev, stages recomputed), saves, affect! on an
// upcrossing / affect_neg! on a downcrossing (a missing one modifies nothing).
// Returns terminate!.
template <int CB, int N, typename Lane, typename SaveFn>
__device__ bool ccb_apply(Lane& L, const Params& P, i64 j, double cb_t, double bs, const double* ev,
                          SaveFn& savevalues, bool& saved_in_cb) {
    typedef OdemiCcb<CB> C;
    Ctl& c = L.c;
    c.dtpropose = c.tdir * jl_max(nextfloat_pos(P.dtmin), c.tdir * c.dt);
    if (cb_t != c.t) {
        double v[N];
        lane_interp(L, (cb_t - c.tprev) / c.dt, v);
#pragma unroll
        for (int i = 0; i < N; i++) L.u[i] = v[i];
        c.t = cb_t;
        c.dt = c.t - c.tprev;
        lane_addsteps(L, P);
        c.reeval_fsal = 1;
    }
    const bool savedexactly = savevalues();
    saved_in_cb = true;
    if (C::save0 && !savedexactly) save_point<N>(c, P, j, c.t, L.u);
    bool modified = true, term = false;
    if constexpr (C::vlen > 0) {
        // VectorContinuousCallback: affect!(integrator, simultaneous_events)
        if constexpr (C::has_affect != 0) term = C::template vaffect<double>(L.u, L.p, c.t, ev);
        else modified = false;
    } else if (bs < 0) {
        if constexpr (C::has_affect != 0) { C::template affect<double>(L.u, L.p, c.t); term = C::term != 0; }
        else modified = false;
    } else if (bs > 0) {
        if constexpr (C::has_affect_neg != 0) {
            C::template affect_neg<double>(L.u, L.p, c.t);
            term = C::term_neg != 0;
        } else {
            modified = false;
        }
    }
    if (modified) {
        c.reeval_fsal = 1;
        if (C::save1) { save_point<N>(c, P, j, c.t, L.u); saved_in_cb = true; }
    }
    return term;
}

#ifndef ODEMI_VCB_MAXLEN
#define ODEMI_VCB_MAXLEN 1
#endif
// get_condition for a VectorContinuousCallback: all vlen values at abst
template <int CB, int N, typename Lane>
__device__ __forceinline__ void vcb_condition(Lane& L, double abst, double* out) {
    Ctl& c = L.c;
    c.ncond++;
    if (abst == c.t) { OdemiCcb<CB>::template vcond<double>(out, L.u, L.p, abst); return; }
    if (abst == c.tprev) { OdemiCcb<CB>::template vcond<double>(out, L.uprev, L.p, abst); return; }
    double v[N];
    lane_interp(L, (abst - c.tprev) / c.dt, v);
    OdemiCcb<CB>::template vcond<double>(out, v, L.p, abst);
}
// findall_events! (callbacks.jl:499-521): event i when the bottom sign is
// nonzero and bottom sign * value <= 0
template <int V>
__device__ __forceinline__ unsigned vcb_events(const double* bs, const double* val) {
    unsigned m = 0u;
#pragma unroll
    for (int i = 0; i < V; i++)
        if (bs[i] != 0 && bs[i] * val[i] <= 0) m |= 1u << i;
    return m;
}
// find_callback_time for a VectorContinuousCallback (callbacks.jl:239-362):
// after this callback's own last event the bottom signs of the components
// that fired are re-read at the nudged tprev (nudge_tprev on the value
// closest to zero among them); the interp_points checks run unless every
// component has an event at t; each event component is rooted to adjacent
// floats (from the nudged tprev when it fired last time), the earliest root
// wins, and components rooted at that same time are simultaneous, with
// direction -sign(bottom), written to ev.
template <int CB, int N, typename Lane>
__device__ bool vcb_find_time(Lane& L, double& cb_t, double* ev, double& residual) {
    typedef OdemiCcb<CB> C;
    constexpr int V = C::vlen;
    Ctl& c = L.c;
    double bc[V], bs[V], top[V], tmp[V];
    const double bottom_t = c.tprev;
    vcb_condition<CB, N>(L, bottom_t, bc);
#pragma unroll
    for (int i = 0; i < V; i++) { bs[i] = jl_sign(bc[i]); ev[i] = 0.0; }
    double nudged_t = bottom_t;
    const unsigned prev = (c.event_last == CB + 1) ? c.sim_mask : 0u;
    if (c.event_last == CB + 1) {
        double minval = 0.0, minabs = __builtin_inf();
#pragma unroll
        for (int i = 0; i < V; i++)
            if (((prev >> i) & 1u) && fabs(bc[i]) < minabs) { minabs = fabs(bc[i]); minval = bc[i]; }
        if (fabs(minval - c.last_event_error) <= C::abstol) nudged_t = c.tprev + c.dt * C::repeat_nudge;
        vcb_condition<CB, N>(L, nudged_t, tmp);
#pragma unroll
        for (int i = 0; i < V; i++)
            if ((prev >> i) & 1u) bs[i] = jl_sign(tmp[i]);
    }
    double top_t = c.t;
    vcb_condition<CB, N>(L, top_t, top);
    unsigned mask = vcb_events<V>(bs, top);
    constexpr unsigned all = (V >= 32) ? 0xffffffffu : ((1u << V) - 1u);
    if (C::interp_points != 0 && mask != all) {
        const double h = (c.t - c.tprev) / (double)(C::interp_points - 1);
        for (int i = 2; i <= C::interp_points; i++) {
            top_t = (i == C::interp_points) ? c.t : fma((double)(i - 1), h, c.tprev);
            vcb_condition<CB, N>(L, top_t, top);
            mask = vcb_events<V>(bs, top);
            if (mask) break;
        }
    }
    if (!mask) { cb_t = c.t; residual = 0.0; return false; }
    if (C::rootfind == 2) {
#pragma unroll
        for (int i = 0; i < V; i++)
            if ((mask >> i) & 1u) ev[i] = -bs[i];
        cb_t = top_t; residual = 0.0;
        return true;
    }
    // rightfloat(top_t, tdir): nextfloat forward, prevfloat backward
    double cbt = (L.c.tdir > 0) ? __longlong_as_double(__double_as_longlong(top_t) + (top_t >= 0.0 ? 1 : -1))
                                : __longlong_as_double(__double_as_longlong(top_t) + (top_t > 0.0 ? -1 : 1)), res = 0.0;
    if (top_t == 0.0) cbt = (L.c.tdir > 0) ? __longlong_as_double(1) : -__longlong_as_double(1);
#pragma unroll
    for (int idx = 0; idx < V; idx++) {
        if (!((mask >> idx) & 1u)) continue;
        double cbi;
        if (jl_sign(top[idx]) == 0.0) {
            cbi = top_t;
        } else {
            double l = ((prev >> idx) & 1u) ? nudged_t : bottom_t, r = top_t;
            for (int it = 0; it < 2200; it++) {
                const double m = l + 0.5 * (r - l);
                if (m == l || m == r) break;
                vcb_condition<CB, N>(L, m, tmp);
                const double sm = jl_sign(tmp[idx]);
                if (sm == 0.0) { l = m; r = m; break; }
                if (sm == bs[idx]) l = m;
                else r = m;
            }
            cbi = (C::rootfind == 1) ? r : l;
        }
        if (L.c.tdir * cbi < L.c.tdir * cbt) {
#pragma unroll
            for (int i = 0; i < V; i++) ev[i] = 0.0;
        }
        if (L.c.tdir * cbi <= L.c.tdir * cbt) {
            cbt = cbi;
            vcb_condition<CB, N>(L, cbi, tmp);
            res = tmp[idx];
            ev[idx] = -bs[idx];
        }
    }
    cb_t = cbt; residual = res;
    return true;
}

template <int CB, int N, typename Lane>
__device__ __forceinline__ void ccb_find_first(Lane& L, int& win, double& tmin, double& bs_win, double& res_win,
                                               double* ev_win) {
    if constexpr (CB < ODEMI_NCCB) {
        typedef OdemiCcb<CB> C;
        double cbt, bs = 0.0, rs;
        if constexpr (C::vlen > 0) {
            double ev[C::vlen];
            if (vcb_find_time<CB, N>(L, cbt, ev, rs) && (win < 0 || L.c.tdir * cbt < L.c.tdir * tmin)) {
                win = CB; tmin = cbt; bs_win = bs; res_win = rs;
#pragma unroll
                for (int i = 0; i < C::vlen; i++) ev_win[i] = ev[i];
            }
        } else {
            if (ccb_find_time<CB, N>(L, cbt, bs, rs) && (win < 0 || L.c.tdir * cbt < L.c.tdir * tmin)) {
                win = CB; tmin = cbt; bs_win = bs; res_win = rs;
#pragma unroll
                for (int i = 0; i < ODEMI_VCB_MAXLEN; i++) ev_win[i] = 0.0;
            }
        }
        ccb_find_first<CB + 1, N>(L, win, tmin, bs_win, res_win, ev_win);
    }
}
template <int CB, int N, typename Lane, typename SaveFn>
__device__ __forceinline__ bool ccb_apply_winner(Lane& L, const Params& P, i64 j, int win, double cbt, double bs,
                                                 const double* ev, SaveFn& savevalues, bool& saved_in_cb) {
    if constexpr (CB < ODEMI_NCCB) {
        if (win == CB) return ccb_apply<CB, N>(L, P, j, cbt, bs, ev, savevalues, saved_in_cb);
        return ccb_apply_winner<CB + 1, N>(L, P, j, win, cbt, bs, ev, savevalues, saved_in_cb);
    } else {
        return false;
    }
}
// find_first_continuous_callback (callbacks.jl:157-237): the earliest event
// wins, the first callback on ties; then apply_callback! for it.
template <int N, typename Lane, typename SaveFn>
__device__ bool lane_continuous_callbacks(Lane& L, const Params& P, i64 j, SaveFn& savevalues, bool& saved_in_cb) {
    int win = -1;
    double tmin = 0.0, bs = 0.0, rs = 0.0;
    double ev[ODEMI_VCB_MAXLEN];
#pragma unroll
    for (int i = 0; i < ODEMI_VCB_MAXLEN; i++) ev[i] = 0.0;
    ccb_find_first<0, N>(L, win, tmin, bs, rs, ev);
    if (win < 0) { L.c.event_last = 0; return false; }
    L.c.last_event_error = rs;
    L.c.event_last = win + 1;
    unsigned m = 0u;
#pragma unroll
    for (int i = 0; i < ODEMI_VCB_MAXLEN; i++)
        if (ev[i] != 0.0) m |= 1u << i;
    L.c.sim_mask = m;
    return ccb_apply_winner<0, N>(L, P, j, win, tmin, bs, ev, savevalues, saved_in_cb);
}
#endif

// One step attempt: loopheader! + check_error! + perform_step! + loopfooter!.
// Returns 0 while running, otherwise the final retcode (lane finished).
template <bool SAVES, typename S, typename K, int N>
__device__ __forceinline__ int tsit5_attempt(Tsit5Lane<S, K, N>& L, const Params& P, i64 j) {
    const CtlConst kc = ctl_const(false);
    Ctl& c = L.c;
    if (!(c.tdir * c.t < c.tdir * c.tf)) return RC_SUCCESS;
    if (c.iter > 0) {
        if (c.accept_step) {
            c.success_iter++;
#pragma unroll
            for (int i = 0; i < N; i++) { L.uprev[i] = L.u[i]; L.k1[i] = L.k7[i]; }
#if defined(ODEMI_NDCB) || defined(ODEMI_NCCB)
            if (c.reeval_fsal) {   // update_fsal! -> reset_fsal! (integrator_utils.jl:215-239,1325-1338)
                K uk[N];
#pragma unroll
                for (int i = 0; i < N; i++) uk[i] = (K)L.u[i];
                odemi_f<K>(L.k1, uk, L.p, (K)c.t);
                c.nf += 1;
                c.reeval_fsal = 0;
            }
#endif
            c.dt = c.dtpropose;
            apply_step_tstop_clip(c);
        } else {
            if (c.isout) c.dt = c.dt * kc.qmin;   // handle_step_rejection! (integrator_utils.jl:135-136)
            else reject_controller<(sizeof(K) < sizeof(S))>(c, kc);
        }
    }
    c.iter++;
    fix_dt_at_bounds(c, P.dtmin);
    modify_dt_for_tstops(c);
    int rc = check_error(c, P, all_finite<S, N>(L.u));
    if (rc != RC_SUCCESS) return rc;
    if (c.next_step_tstop && fabs(c.dt) < jl_eps(fabs(c.t))) {
        c.accept_step = 1;
    } else {
        tsit5_perform_step(L, P);
    }
    double ttmp = c.t + c.dt;
    constexpr bool HW = sizeof(K) < sizeof(S);   // mixed: hardware FastPower, Newton-refined FP64 quotients
    double q = stepsize_controller<HW>(c, kc);
#ifdef ODEMI_HAS_ISOUT
    c.isout = odemi_isout<double, K>(L.u, L.p, ttmp) ? 1 : 0;   // integrator_utils.jl:612-614
#endif
    c.accept_step = !c.isout && (c.EEst <= 1.0);
    if (c.accept_step) {
        c.naccept++;
        c.tprev = c.t;
        if (c.next_step_tstop) { c.dt = c.dtpropose; c.next_step_tstop = 0; c.t = c.tnext; }
        else c.t = ttmp;
        accept_controller<HW>(c, kc, q, P.dtmin);
        auto savefn = [&]() -> bool {
            return (SAVES && (P.nsaveat > 0 || P.save_everystep)) ? tsit5_savevalues(L, P, j) : false;
        };
        const bool term = handle_callbacks<N>(c, P, j, L.u, L.p, savefn, [&](bool& saved_in_cb) -> bool {
#ifdef ODEMI_NCCB
            return lane_continuous_callbacks<N>(L, P, j, savefn, saved_in_cb);
#else
            (void)saved_in_cb;
            return false;
#endif
        });
        handle_tstop(c, P);   // after the callbacks: an event can move t back before the tstop
        if (term) return RC_TERMINATED;
    } else {
        c.nreject++;
    }
    return (c.tdir * c.t < c.tdir * c.tf) ? 0 : RC_SUCCESS;
}

// ---------------------------------------------------------------------------
// Work queue.  A wave takes queue slots from the global head in chunks of
// ODEMI_QUEUE_CHUNK (>= 64) with one atomic and deals them to its idle lanes
// as they free up.  One device-wide counter shared by every wave of all XCDs
// is a serialisation point: with one atomic per refill (a lane finishes every
// few step attempts of a wave) it bounded the whole lane engine; chunking cuts
// the atomics by the chunk size.  Slots beyond the end come back >= N.
// `q` is wave-uniform; every lane of the wave calls this together.
// ---------------------------------------------------------------------------
#ifndef ODEMI_QUEUE_CHUNK
#define ODEMI_QUEUE_CHUNK 64
#endif
static_assert(ODEMI_QUEUE_CHUNK >= 64, "a refill must cover a whole wave");
struct WaveQueue {
    u64 next = 0, end = 0;   // slots [next, end) taken from the head, not yet dealt
};
__device__ __forceinline__ i64 wave_fetch(u64* head, WaveQueue& q, bool need) {
    const u64 mask = __ballot(need);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const u64 cnt = (u64)__popcll(mask);
    const u64 rank = (u64)__popcll(mask & ((lane == 0) ? 0ULL : (~0ULL >> (64 - lane))));
    const u64 avail = q.end - q.next;
    i64 slot = (i64)(q.next + rank);
    if (avail < cnt) {
        // the undealt rest first, then a fresh chunk
        const int leader = __ffsll((long long)mask) - 1;
        u64 base = 0;
        if (lane == leader) base = atomicAdd(head, (u64)ODEMI_QUEUE_CHUNK);
        base = __shfl(base, leader);
        if (rank >= avail) slot = (i64)(base + (rank - avail));
        q.next = base + (cnt - avail);
        q.end = base + (u64)ODEMI_QUEUE_CHUNK;
    } else {
        q.next += cnt;
    }
    return need ? slot : -1;
}

// Refill in batches: idle lanes fetch only once ODEMI_REFILL_MIN of them are
// idle (or no lane is active), so the start-up code runs for several lanes at
// once instead of once per finished lane with the rest of the wave waiting.
// Active lanes always finish, so waiting lanes are reached.  16 measured best
// of 1/4/8/16 on the Lorenz ensemble (-1.2 % kernel time; Robertson flat).
#ifndef ODEMI_REFILL_MIN
#define ODEMI_REFILL_MIN 16
#endif
__device__ __forceinline__ bool refill_now(bool need, bool active) {
    const u64 needm = __ballot(need);
    return need && (__popcll(needm) >= ODEMI_REFILL_MIN || __ballot(active) == 0);
}

template <bool SAVES, typename S, typename K, int N>
__device__ void tsit5_persistent(const Params& P) {
    Tsit5Lane<S, K, N> L;
    WaveQueue q;
    bool active = false, exhausted = false;
    i64 j = 0, js = 0;
    for (;;) {
        bool need = refill_now(!active && !exhausted, active);
        i64 slot = wave_fetch(&P.counters[0], q, need);
        if (need) {
            if (slot < P.N) {
                js = slot;
                j = P.order ? P.order[slot] : slot;
                tsit5_init(L, P, j, slot);
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            int rc = tsit5_attempt<SAVES>(L, P, j);
            if (rc) {
                if (!SAVES && P.rec_out) finish_lane_rec<N>(L.c, P, js, L.u, rc);
                else finish_lane<N>(L.c, P, j, L.u, rc);
                active = false;
            }
        }
    }
}

// Plain engine: one kernel launch per step attempt, state in HBM.
// st_scalar rows: t dt dtpropose tprev last_event_error tf dtmax EEst q11 errold
// st_int rows: iter success_iter accept_step next_step_tstop nf naccept nreject done
//               saveat cursor, njacs nw nsolve (Rodas5P), tstops cursor,
//               isout | reeval_fsal << 1, ncondition, event_last | sim_mask << 8
// st_vec rows: u uprev k1 k7 (Tsit5)
template <int N>
__device__ __forceinline__ void plain_load_ctl(Ctl& c, const Params& P, i64 j) {
    const double* s = P.st_scalar;
    const i64 M = P.N;
    c.t = s[0 * M + j]; c.dt = s[1 * M + j]; c.dtpropose = s[2 * M + j]; c.tprev = s[3 * M + j];
    c.last_event_error = s[4 * M + j];
    c.tf = s[5 * M + j]; c.dtmax = s[6 * M + j];
    c.tdir = (c.tf < P.t0[j]) ? -1.0 : 1.0;
    c.EEst = s[7 * M + j]; c.q11 = s[8 * M + j]; c.errold = s[9 * M + j];
    const int* q = P.st_int;
    c.iter = q[0 * M + j]; c.success_iter = q[1 * M + j]; c.accept_step = q[2 * M + j];
    c.next_step_tstop = q[3 * M + j]; c.nf = q[4 * M + j]; c.naccept = q[5 * M + j];
    c.nreject = q[6 * M + j];
    c.nsaved = P.nsaved ? P.nsaved[j] : 0;
    c.spos = 0; c.send = 0; c.tf_in_saveat = 0;
    if (P.nsaveat > 0) {
        saveat_init(c, P, P.t0[j]);
        c.spos = q[8 * M + j];
    }
    c.tsp = 0; c.tsend = 0; c.tnext = c.tf;
    if (P.ntstops > 0) {
        tstops_init(c, P, P.t0[j]);
        c.tsp = q[12 * M + j];
        c.tnext = (c.tsp != c.tsend) ? P.tstops[c.tsp] : c.tf;
    }
    const int fl = q[13 * M + j];
    c.isout = fl & 1; c.reeval_fsal = (fl >> 1) & 1;
    c.ncond = q[14 * M + j];
    c.event_last = q[15 * M + j] & 0xff;
    c.sim_mask = (unsigned)q[15 * M + j] >> 8;
    c.njacs = 0; c.nw = 0; c.nsolve = 0;
    c.erroldp = __builtin_nanf(""); c.lgEEst = 0.0f;
}
template <int N>
__device__ __forceinline__ void plain_store_ctl(const Ctl& c, const Params& P, i64 j, int done) {
    double* s = P.st_scalar;
    const i64 M = P.N;
    s[0 * M + j] = c.t; s[1 * M + j] = c.dt; s[2 * M + j] = c.dtpropose; s[3 * M + j] = c.tprev;
    s[4 * M + j] = c.last_event_error;
    s[5 * M + j] = c.tf; s[6 * M + j] = c.dtmax;
    s[7 * M + j] = c.EEst; s[8 * M + j] = c.q11; s[9 * M + j] = c.errold;
    int* q = P.st_int;
    q[0 * M + j] = c.iter; q[1 * M + j] = c.success_iter; q[2 * M + j] = c.accept_step;
    q[3 * M + j] = c.next_step_tstop; q[4 * M + j] = c.nf; q[5 * M + j] = c.naccept;
    q[6 * M + j] = c.nreject; q[7 * M + j] = done; q[8 * M + j] = c.spos; q[12 * M + j] = c.tsp;
    q[13 * M + j] = c.isout | (c.reeval_fsal << 1);
    q[14 * M + j] = c.ncond;
    q[15 * M + j] = (int)((unsigned)c.event_last | (c.sim_mask << 8));
    if (P.nsaved) P.nsaved[j] = c.nsaved;
}

template <int N>
__device__ void tsit5_plain_init(const Params& P) {
    i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    Tsit5Lane<double, double, N> L;
    tsit5_init(L, P, j);
    plain_store_ctl<N>(L.c, P, j, 0);
    const i64 M = P.N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        P.st_vec[((i64)0 * P.nmax + i) * M + j] = L.u[i];
        P.st_vec[((i64)1 * P.nmax + i) * M + j] = L.uprev[i];
        P.st_vec[((i64)2 * P.nmax + i) * M + j] = L.k1[i];
        P.st_vec[((i64)3 * P.nmax + i) * M + j] = L.k7[i];
    }
}

template <int N>
__device__ void tsit5_plain_step(const Params& P) {
    i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    if (P.st_int[7 * P.N + j]) return;
    Tsit5Lane<double, double, N> L;
    plain_load_ctl<N>(L.c, P, j);
    const i64 M = P.N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        L.u[i] = P.st_vec[((i64)0 * P.nmax + i) * M + j];
        L.uprev[i] = P.st_vec[((i64)1 * P.nmax + i) * M + j];
        L.k1[i] = P.st_vec[((i64)2 * P.nmax + i) * M + j];
        L.k7[i] = P.st_vec[((i64)3 * P.nmax + i) * M + j];
    }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
    int rc = tsit5_attempt<true>(L, P, j);
    if (rc) {
        finish_lane<N>(L.c, P, j, L.u, rc);
        plain_store_ctl<N>(L.c, P, j, 1);
    } else {
        plain_store_ctl<N>(L.c, P, j, 0);
#pragma unroll
        for (int i = 0; i < N; i++) {
            P.st_vec[((i64)0 * P.nmax + i) * M + j] = L.u[i];
            P.st_vec[((i64)1 * P.nmax + i) * M + j] = L.uprev[i];
            P.st_vec[((i64)2 * P.nmax + i) * M + j] = L.k1[i];
            P.st_vec[((i64)3 * P.nmax + i) * M + j] = L.k7[i];
        }
        atomicAdd(&P.counters[1], 1ULL);
    }
}

// ---------------------------------------------------------------------------
// Dense LU on a lane-resident matrix (compile-time N, branchless row swaps).
// Textbook right-looking LU with partial pivoting (first maximum on ties),
// multipliers scaled by the reciprocal pivot, trailing update a -= l*u as fma,
// triangular solves in LAPACK's column order.  In FP32 (the factorisation that
// FP64 refinement corrects) the pivot reciprocal and the back substitution use
// the hardware reciprocal instead of IEEE division.
// ---------------------------------------------------------------------------
template <typename T, int N>
__device__ __forceinline__ void lu_factor_reg(T (&A)[N][N], int (&piv)[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        int p = k;
        T amax = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            T v = fabs(A[i][k]);
            if (v > amax) { amax = v; p = i; }
        }
        piv[k] = p;
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            bool s = (p == i);
#pragma unroll
            for (int c = 0; c < N; c++) {
                T a = A[k][c], b = A[i][c];
                A[k][c] = s ? b : a;
                A[i][c] = s ? a : b;
            }
        }
        T inv;
        if constexpr (sizeof(T) == sizeof(float)) inv = __builtin_amdgcn_rcpf(A[k][k]);
        else inv = (T)1 / A[k][k];
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            T l = A[i][k] * inv;
            A[i][k] = l;
#pragma unroll
            for (int c = k + 1; c < N; c++) A[i][c] = fma(-l, A[k][c], A[i][c]);
        }
    }
}
template <typename T, int N>
__device__ __forceinline__ void lu_solve_reg(const T (&A)[N][N], const int (&piv)[N], T (&b)[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            bool s = (piv[k] == i);
            T x = b[k], y = b[i];
            b[k] = s ? y : x;
            b[i] = s ? x : y;
        }
    }
#pragma unroll
    for (int i = 1; i < N; i++) {
        T acc = b[i];
#pragma unroll
        for (int c = 0; c < i; c++) acc = fma(-A[i][c], b[c], acc);
        b[i] = acc;
    }
    // back substitution in column order (LAPACK dtrsv 'U','N','N')
#pragma unroll
    for (int c = N - 1; c >= 0; c--) {
        if constexpr (sizeof(T) == sizeof(float)) b[c] = b[c] * __builtin_amdgcn_rcpf(A[c][c]);
        else b[c] = b[c] / A[c][c];
#pragma unroll
        for (int i = 0; i < c; i++) b[i] = fma(-A[i][c], b[c], b[i]);
    }
}

// ---------------------------------------------------------------------------
// Rodas5P, one trajectory per lane (small n).  FP64 state, RHS and stages.
//   F = factorisation type: double (FP64 LU) or float (FP32 LU with FP64
//   iterative refinement against the FP64 W, P.refine_sweeps sweeps).
// ---------------------------------------------------------------------------
template <typename F, int N>
struct RodasLane {
    Ctl c;
    double u[N], uprev[N];
    double J[N][N];
    double dT[N];
    double ks[8][N];
    double dense[3][N];
    double p[ODEMI_NP];
};

template <typename F, int N>
__device__ __forceinline__ void rodas_init(RodasLane<F, N>& L, const Params& P, i64 j) {
    ctl_init(L.c, P, j);
#pragma unroll
    for (int i = 0; i < N; i++) { L.u[i] = P.u0[(i64)i * P.N + j]; L.uprev[i] = L.u[i]; L.dT[i] = 0.0; }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
        for (int b = 0; b < N; b++) L.J[a][b] = 0.0;
    if (P.save_start) save_point<N>(L.c, P, j, L.c.t, L.u);
    if (P.pre_dt) {  // computed by rodas_prep with the same arithmetic
        L.c.dt = P.pre_dt[j];
        if (P.dt_user == 0.0) L.c.nf += 2;
    } else if (P.dt_user != 0.0) {  // user dt (a NaN dt is reported as DtNaN)
        L.c.dt = user_dt(L.c, P);
    } else {
        L.c.dt = initdt_lane<N>(L.c, L.u, L.p, P.reltol, P.abstol, L.c.dtmax, P.dtmin);
        L.c.nf += 2;
    }
    L.c.dtpropose = L.c.dt;
}

// initdt for every trajectory ahead of the persistent kernel (see tsit5_prep).
template <int N>
__device__ void rodas_prep(const Params& P) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    Ctl c;
    ctl_init(c, P, j);
    double u[N], p[ODEMI_NP];
#pragma unroll
    for (int i = 0; i < N; i++) u[i] = P.u0[(i64)i * P.N + j];
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) p[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
    P.pre_dt[j] = (P.dt_user != 0.0) ? user_dt(c, P) : initdt_lane<N>(c, u, p, P.reltol, P.abstol, c.dtmax, P.dtmin);
}

// W x = b with W held in FP64; returns x in b.
template <typename F, int N>
__device__ __forceinline__ void rodas_wsolve(const double (&W)[N][N], const F (&LU)[N][N], const int (&piv)[N],
                                             double (&b)[N], int sweeps) {
    if constexpr (sizeof(F) == sizeof(double)) {
        lu_solve_reg<double, N>(LU, piv, b);
    } else {
        F xf[N];
        double x[N];
#pragma unroll
        for (int i = 0; i < N; i++) xf[i] = (F)b[i];
        lu_solve_reg<F, N>(LU, piv, xf);
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = (double)xf[i];
        for (int s = 0; s < sweeps; s++) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                double acc = b[i];
#pragma unroll
                for (int c = 0; c < N; c++) acc = fma(-W[i][c], x[c], acc);
                xf[i] = (F)acc;
            }
            lu_solve_reg<F, N>(LU, piv, xf);
#pragma unroll
            for (int i = 0; i < N; i++) x[i] = x[i] + (double)xf[i];
        }
#pragma unroll
        for (int i = 0; i < N; i++) b[i] = x[i];
    }
}

template <typename F, int N>
__device__ __forceinline__ void rodas_perform_step(RodasLane<F, N>& L, const Params& P) {
    constexpr bool HW = sizeof(F) < sizeof(double);   // mixed: Newton-refined FP64 reciprocals and norm
    Ctl& c = L.c;
    const double t = c.t, dt = c.dt;
    const double idt = HW ? rcp_nr(dt) : 1.0 / dt;
    const double dtgamma = dt * RodasTab::gamma;
    double fsal[N], du[N], lt[N];
    odemi_f<double>(fsal, L.uprev, L.p, t);
    c.nf += 1;
    // do_newJW (derivative_utils.jl:535-595): J is reused only right after a reject
    bool new_jac = (c.iter <= 1) ? true : !(c.iter > 1 && c.EEst > 1.0);
    if (new_jac) {
        odemi_jac<double>(&L.J[0][0], L.uprev, L.p, t);
        c.njacs += 1;
    }
    // jacobian2W! : W = J - I/(dt*gamma)
    const double invdtgamma = HW ? rcp_nr(dtgamma) : 1.0 / dtgamma;
    double W[N][N];
    F LU[N][N];
    int piv[N];
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
        for (int b = 0; b < N; b++) W[a][b] = (a == b) ? fma(-1.0, invdtgamma, L.J[a][b]) : L.J[a][b];
    c.nw += 1;
    if (new_jac) {
        odemi_tgrad<double>(L.dT, L.uprev, L.p, t);
        if (!ODEMI_HAS_TGRAD) c.nf += 1;
    }
    const double dtd0 = dt * kRd[0];
#pragma unroll
    for (int i = 0; i < N; i++) lt[i] = fsal[i] + dtd0 * L.dT[i];
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
        for (int b = 0; b < N; b++) LU[a][b] = (F)W[a][b];
    lu_factor_reg<F, N>(LU, piv);
    rodas_wsolve<F, N>(W, LU, piv, lt, P.refine_sweeps);
#pragma unroll
    for (int i = 0; i < N; i++) L.ks[0][i] = -lt[i];
    c.nsolve += 1;
#pragma unroll
    for (int s = 1; s < 8; s++) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = L.uprev[i];
#pragma unroll
            for (int q = 0; q < s; q++) acc = fma(kRA[s][q], L.ks[q][i], acc);
            L.u[i] = acc;
        }
        odemi_f<double>(du, L.u, L.p, fma(kRc[s], dt, t));
        c.nf += 1;
        const double dtds = dt * kRd[s];
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = fma(dtds, L.dT[i], du[i]);
#pragma unroll
            for (int q = 0; q < s; q++) acc = fma(kRC[s][q] * idt, L.ks[q][i], acc);
            lt[i] = acc;
        }
        rodas_wsolve<F, N>(W, LU, piv, lt, P.refine_sweeps);
#pragma unroll
        for (int i = 0; i < N; i++) L.ks[s][i] = -lt[i];
        c.nsolve += 1;
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        double acc = L.uprev[i];
#pragma unroll
        for (int q = 0; q < 7; q++) acc = fma(kRA[7][q], L.ks[q][i], acc);
        acc = fma(1.0, L.ks[7][i], acc);
        L.u[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 7; q++) acc = fma(0.0, L.ks[q][i], acc);
        acc = fma(1.0, L.ks[7][i], acc);
        du[i] = acc;
    }
    if constexpr (HW) c.EEst = residual_norm_nr<N>(du, L.uprev, L.u, P.abstol, P.reltol);
    else c.EEst = residual_norm<N>(du, L.uprev, L.u, P.abstol, P.reltol);
#ifdef ODEMI_NCCB
    constexpr bool dense_always = true;   // event search interpolates every step
#else
    constexpr bool dense_always = false;
#endif
    if (dense_always || P.nsaveat > 0) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int i = 0; i < N; i++) {
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < 8; q++) acc = fma(kRH[r][q], L.ks[q][i], acc);
                L.dense[r][i] = acc;
            }
    }
}

// rosenbrock_interpolants.jl:172-206 (interp_order 3)
template <typename F, int N>
__device__ __forceinline__ void rodas_interp(const RodasLane<F, N>& L, double th, double* v) {
    const double th1 = 1.0 - th;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double in1 = fma(th, L.dense[2][i], L.dense[1][i]);
        double in2 = fma(th, in1, L.dense[0][i]);
        double in3 = fma(th1, in2, L.u[i]);
        v[i] = fma(th, in3, th1 * L.uprev[i]);
    }
}

template <typename F, int N>
__device__ __forceinline__ bool rodas_savevalues(RodasLane<F, N>& L, const Params& P, i64 j) {
    Ctl& c = L.c;
    bool exact = false;
    while (c.spos != c.send && c.tdir * P.saveat[c.spos] <= c.tdir * c.t) {
        double curt = P.saveat[c.spos];
        c.spos += (c.tdir > 0) ? 1 : -1;
        if (curt != c.t) {
            double v[N];
            rodas_interp(L, (curt - c.tprev) / c.dt, v);
            save_point<N>(c, P, j, curt, v);
        } else {
            if (curt == c.tf && !P.save_end) continue;
            save_point<N>(c, P, j, c.t, L.u);
            exact = true;
        }
    }
    if (P.save_everystep) {
        int s = c.nsaved;
        double last = (s > 0 && s - 1 < P.max_saves) ? P.t_out[(i64)(s - 1) * P.N + j] : __longlong_as_double(0x7ff8000000000000LL);
        if (s == 0 || ((c.t != last || c.dt == 0.0) && (P.save_end || c.t != c.tf))) {
            save_point<N>(c, P, j, c.t, L.u);
            exact = true;
        }
    }
    return exact;
}

#ifdef ODEMI_NCCB
// _ode_addsteps! for RosenbrockCache (stiff_addsteps.jl:140-224) with
// always_calc_begin, after change_t_via_interpolation!: the stages of the step
// from uprev at tprev with the new dt, reusing the step's J and dT and its
// fsalfirst f(uprev, tprev); W refactored; no @muladd in that function, so
// only linsolve_tmp = fsalfirst + dtgamma dT is fused; dtC = C ./ dt; the
// dense rows from the new stages.  No stats change.
template <typename F, int N>
__device__ void rodas_addsteps(RodasLane<F, N>& L, const Params& P) {
    constexpr bool HW = sizeof(F) < sizeof(double);
    const double t = L.c.tprev, dt = L.c.dt;
    const double dtgamma = dt * RodasTab::gamma;
    double fsal[N], du[N], lt[N], st[N];
    odemi_f<double>(fsal, L.uprev, L.p, t);
#pragma unroll
    for (int i = 0; i < N; i++) lt[i] = fma(dtgamma, L.dT[i], fsal[i]);
    const double invdtgamma = HW ? rcp_nr(dtgamma) : 1.0 / dtgamma;
    double W[N][N];
    F LU[N][N];
    int piv[N];
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
        for (int b = 0; b < N; b++) {
            W[a][b] = (a == b) ? fma(-1.0, invdtgamma, L.J[a][b]) : L.J[a][b];
            LU[a][b] = (F)W[a][b];
        }
    lu_factor_reg<F, N>(LU, piv);
    rodas_wsolve<F, N>(W, LU, piv, lt, P.refine_sweeps);
#pragma unroll
    for (int i = 0; i < N; i++) L.ks[0][i] = -lt[i];
#pragma unroll
    for (int s = 1; s < 8; s++) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = L.uprev[i];
#pragma unroll
            for (int q = 0; q < s; q++) acc = acc + kRA[s][q] * L.ks[q][i];
            st[i] = acc;
        }
        odemi_f<double>(du, st, L.p, t + kRc[s] * dt);
        const double dtds = dt * kRd[s];
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = du[i] + dtds * L.dT[i];
#pragma unroll
            for (int q = 0; q < s; q++) acc = acc + (kRC[s][q] / dt) * L.ks[q][i];
            lt[i] = acc;
        }
        rodas_wsolve<F, N>(W, LU, piv, lt, P.refine_sweeps);
#pragma unroll
        for (int i = 0; i < N; i++) L.ks[s][i] = -lt[i];
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) acc = acc + kRH[r][q] * L.ks[q][i];
            L.dense[r][i] = acc;
        }
}
template <typename F, int N>
__device__ __forceinline__ void lane_interp(const RodasLane<F, N>& L, double th, double* v) {
    rodas_interp(L, th, v);
}
template <typename F, int N>
__device__ __forceinline__ void lane_addsteps(RodasLane<F, N>& L, const Params& P) {
    rodas_addsteps(L, P);
}
#endif

template <typename F, int N>
__device__ __forceinline__ int rodas_attempt(RodasLane<F, N>& L, const Params& P, i64 j) {
    const CtlConst kc = ctl_const(true);
    Ctl& c = L.c;
    if (!(c.tdir * c.t < c.tdir * c.tf)) return RC_SUCCESS;
    if (c.iter > 0) {
        if (c.accept_step) {
            c.success_iter++;
#pragma unroll
            for (int i = 0; i < N; i++) L.uprev[i] = L.u[i];
            c.dt = c.dtpropose;
            apply_step_tstop_clip(c);
        } else {
            if (c.isout) c.dt = c.dt * kc.qmin;   // handle_step_rejection! (integrator_utils.jl:135-136)
            else reject_controller<(sizeof(F) < sizeof(double))>(c, kc);
        }
    }
    c.iter++;
    fix_dt_at_bounds(c, P.dtmin);
    modify_dt_for_tstops(c);
    int rc = check_error(c, P, all_finite<double, N>(L.u));
    if (rc != RC_SUCCESS) return rc;
    if (c.next_step_tstop && fabs(c.dt) < jl_eps(fabs(c.t))) {
        c.accept_step = 1;
    } else {
        rodas_perform_step(L, P);
    }
    double ttmp = c.t + c.dt;
    constexpr bool HW = sizeof(F) < sizeof(double);   // mixed: hardware FastPower, Newton-refined FP64 quotients
    double q = stepsize_controller<HW>(c, kc);
#ifdef ODEMI_HAS_ISOUT
    c.isout = odemi_isout<double, double>(L.u, L.p, ttmp) ? 1 : 0;   // integrator_utils.jl:612-614
#endif
    c.accept_step = !c.isout && (c.EEst <= 1.0);
    if (c.accept_step) {
        c.naccept++;
        c.tprev = c.t;
        if (c.next_step_tstop) { c.dt = c.dtpropose; c.next_step_tstop = 0; c.t = c.tnext; }
        else c.t = ttmp;
        accept_controller<HW>(c, kc, q, P.dtmin);
        auto savefn = [&]() -> bool {
            return (P.nsaveat > 0 || P.save_everystep) ? rodas_savevalues(L, P, j) : false;
        };
        const bool term = handle_callbacks<N>(c, P, j, L.u, L.p, savefn, [&](bool& saved_in_cb) -> bool {
#ifdef ODEMI_NCCB
            return lane_continuous_callbacks<N>(L, P, j, savefn, saved_in_cb);
#else
            (void)saved_in_cb;
            return false;
#endif
        });
        handle_tstop(c, P);
        if (term) return RC_TERMINATED;
    } else {
        c.nreject++;
    }
    return (c.tdir * c.t < c.tdir * c.tf) ? 0 : RC_SUCCESS;
}

template <typename F, int N>
__device__ void rodas_persistent(const Params& P) {
    RodasLane<F, N> L;
    WaveQueue q;
    bool active = false, exhausted = false;
    i64 j = 0;
    for (;;) {
        bool need = refill_now(!active && !exhausted, active);
        i64 slot = wave_fetch(&P.counters[0], q, need);
        if (need) {
            if (slot < P.N) {
                j = P.order ? P.order[slot] : slot;
                rodas_init(L, P, j);
                active = true;
            } else {
                exhausted = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            int rc = rodas_attempt(L, P, j);
            if (rc) {
                finish_lane<N>(L.c, P, j, L.u, rc);
                active = false;
            }
        }
    }
}

// Plain Rodas5P: st_vec rows u, uprev, dT, J[0..N-1]
template <typename F, int N>
__device__ __forceinline__ void rodas_plain_store_vec(const RodasLane<F, N>& L, const Params& P, i64 j) {
    const i64 M = P.N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        P.st_vec[((i64)0 * P.nmax + i) * M + j] = L.u[i];
        P.st_vec[((i64)1 * P.nmax + i) * M + j] = L.uprev[i];
        P.st_vec[((i64)2 * P.nmax + i) * M + j] = L.dT[i];
#pragma unroll
        for (int b = 0; b < N; b++) P.st_vec[((i64)(3 + i) * P.nmax + b) * M + j] = L.J[i][b];
    }
}
template <int N>
__device__ void rodas_plain_init(const Params& P) {
    i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    RodasLane<double, N> L;
    rodas_init(L, P, j);
    plain_store_ctl<N>(L.c, P, j, 0);
    rodas_plain_store_vec(L, P, j);
}
template <int N>
__device__ void rodas_plain_step(const Params& P) {
    i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    if (P.st_int[7 * P.N + j]) return;
    RodasLane<double, N> L;
    plain_load_ctl<N>(L.c, P, j);
    const i64 M = P.N;
    // counters beyond nf/naccept/nreject live in st_int rows 9..11
    L.c.njacs = P.st_int[9 * M + j]; L.c.nw = P.st_int[10 * M + j]; L.c.nsolve = P.st_int[11 * M + j];
#pragma unroll
    for (int i = 0; i < N; i++) {
        L.u[i] = P.st_vec[((i64)0 * P.nmax + i) * M + j];
        L.uprev[i] = P.st_vec[((i64)1 * P.nmax + i) * M + j];
        L.dT[i] = P.st_vec[((i64)2 * P.nmax + i) * M + j];
#pragma unroll
        for (int b = 0; b < N; b++) L.J[i][b] = P.st_vec[((i64)(3 + i) * P.nmax + b) * M + j];
    }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
    int rc = rodas_attempt(L, P, j);
    if (rc) {
        finish_lane<N>(L.c, P, j, L.u, rc);
        plain_store_ctl<N>(L.c, P, j, 1);
    } else {
        plain_store_ctl<N>(L.c, P, j, 0);
        P.st_int[9 * M + j] = L.c.njacs; P.st_int[10 * M + j] = L.c.nw; P.st_int[11 * M + j] = L.c.nsolve;
        rodas_plain_store_vec(L, P, j);
        atomicAdd(&P.counters[1], 1ULL);
    }
}

}  // namespace odemi

namespace odemi {
// ---------------------------------------------------------------------------
// Serving (continuous batching).  The slot pool holds device-resident
// integrator state between ticks in the plain-engine layout (st_scalar,
// st_int, st_vec); st_int row 7 is the slot state (0 free, 1 active).  The
// init kernels admit a contiguous slot range, the tick kernels advance every
// active slot by up to P.tick_attempts step attempts with the state in
// registers.  Stage values are stored as FP64 (exact for FP32 stages), so a
// trajectory's result does not depend on slot placement or tick boundaries.
// ---------------------------------------------------------------------------
template <typename S, typename K, int N>
__device__ __forceinline__ void tsit5_store_state(const Tsit5Lane<S, K, N>& L, const Params& P, i64 j) {
    const i64 M = P.N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        P.st_vec[((i64)0 * P.nmax + i) * M + j] = (double)L.u[i];
        P.st_vec[((i64)1 * P.nmax + i) * M + j] = (double)L.uprev[i];
        P.st_vec[((i64)2 * P.nmax + i) * M + j] = (double)L.k1[i];
        P.st_vec[((i64)3 * P.nmax + i) * M + j] = (double)L.k7[i];
    }
}
// A serving slot's own options: its request's reltol / abstol, written by the
// host into rows 10 / 11 of the pool's scalar state at admission, and (lane
// servers with saves or stopping times) its request's save times, stopping
// times and save flags from the request table.
__device__ __forceinline__ Params slot_params(const Params& P, i64 j) {
    Params Q = P;
    Q.reltol = P.st_scalar[10 * P.N + j];
    Q.abstol = P.st_scalar[11 * P.N + j];
    if (P.req_opts) {
        const int r = P.slot_req[j];
        const int* o = P.req_opts + kReqOpts * (i64)r;
        Q.nsaveat = o[0];
        Q.saveat = P.req_saveat + (i64)r * P.sa_cap;
        Q.ntstops = o[1];
        Q.tstops = P.req_tstops + (i64)r * P.ts_cap;
        Q.save_start = o[2];
        Q.save_end = o[3];
        Q.save_end_user = o[4];
    }
    return Q;
}

template <typename S, typename K, int N>
__device__ void tsit5_serve_init(const Params& P0) {
    i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P0.nslots) return;
    const i64 j = P0.slot0 + q;
    const Params P = slot_params(P0, j);
    Tsit5Lane<S, K, N> L;
    tsit5_init(L, P, j);
    plain_store_ctl<N>(L.c, P, j, 1);
    tsit5_store_state(L, P, j);
}
template <bool SAVES, typename S, typename K, int N>
__device__ void tsit5_serve_tick(const Params& P0) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P0.N) return;
    if (P0.st_int[7 * P0.N + j] != 1) return;
    const Params P = slot_params(P0, j);
    Tsit5Lane<S, K, N> L;
    plain_load_ctl<N>(L.c, P, j);
    const i64 M = P.N;
#pragma unroll
    for (int i = 0; i < N; i++) {
        L.u[i] = (S)P.st_vec[((i64)0 * P.nmax + i) * M + j];
        L.uprev[i] = (S)P.st_vec[((i64)1 * P.nmax + i) * M + j];
        L.k1[i] = (K)P.st_vec[((i64)2 * P.nmax + i) * M + j];
        L.k7[i] = (K)P.st_vec[((i64)3 * P.nmax + i) * M + j];
    }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (K)((k < P.np) ? P.p[(i64)k * P.N + j] : 0.0);
    for (int a = 0; a < P.tick_attempts; a++) {
        int rc = tsit5_attempt<SAVES>(L, P, j);
        if (rc) {
            finish_lane<N>(L.c, P, j, L.u, rc);
            plain_store_ctl<N>(L.c, P, j, 0);
            atomicAdd(&P.req_done[P.slot_req[j]], 1ULL);
            return;
        }
    }
    plain_store_ctl<N>(L.c, P, j, 1);
    tsit5_store_state(L, P, j);
}

template <typename F, int N>
__device__ void rodas_serve_init(const Params& P0) {
    i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P0.nslots) return;
    const i64 j = P0.slot0 + q;
    const Params P = slot_params(P0, j);
    RodasLane<F, N> L;
    rodas_init(L, P, j);
    plain_store_ctl<N>(L.c, P, j, 1);
    const i64 M = P.N;
    P.st_int[9 * M + j] = 0; P.st_int[10 * M + j] = 0; P.st_int[11 * M + j] = 0;
    rodas_plain_store_vec(L, P, j);
}
template <typename F, int N>
__device__ void rodas_serve_tick(const Params& P0) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P0.N) return;
    if (P0.st_int[7 * P0.N + j] != 1) return;
    const Params P = slot_params(P0, j);
    RodasLane<F, N> L;
    plain_load_ctl<N>(L.c, P, j);
    const i64 M = P.N;
    L.c.njacs = P.st_int[9 * M + j]; L.c.nw = P.st_int[10 * M + j]; L.c.nsolve = P.st_int[11 * M + j];
#pragma unroll
    for (int i = 0; i < N; i++) {
        L.u[i] = P.st_vec[((i64)0 * P.nmax + i) * M + j];
        L.uprev[i] = P.st_vec[((i64)1 * P.nmax + i) * M + j];
        L.dT[i] = P.st_vec[((i64)2 * P.nmax + i) * M + j];
#pragma unroll
        for (int b = 0; b < N; b++) L.J[i][b] = P.st_vec[((i64)(3 + i) * P.nmax + b) * M + j];
    }
#pragma unroll
    for (int k = 0; k < ODEMI_NP; k++) L.p[k] = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
    for (int a = 0; a < P.tick_attempts; a++) {
        int rc = rodas_attempt(L, P, j);
        if (rc) {
            finish_lane<N>(L.c, P, j, L.u, rc);
            plain_store_ctl<N>(L.c, P, j, 0);
            atomicAdd(&P.req_done[P.slot_req[j]], 1ULL);
            return;
        }
    }
    plain_store_ctl<N>(L.c, P, j, 1);
    P.st_int[9 * M + j] = L.c.njacs; P.st_int[10 * M + j] = L.c.nw; P.st_int[11 * M + j] = L.c.nsolve;
    rodas_plain_store_vec(L, P, j);
}
}  // namespace odemi

#if defined(ODEMI_WITH_WG) && !defined(ODEMI_RTC)
#include "odemi_wg.h"
#endif
