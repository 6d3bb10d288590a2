// odemi_wg.h -- Rodas5P with one workgroup per trajectory, for systems too
// large for a lane (n up to 128) and for batches of uneven dimension (nvec).
//
// Same step-control semantics as the lane kernels (see odemi_device.h) and the
// same arithmetic order as the CPU restatement: every element of the LU and of
// the triangular solves sees the same fma sequence as the sequential code, and
// the norms are accumulated by one thread in index order, so the FP64 variant
// is bit-identical to the restatement.
//
// LDS layout (dynamic, sized by the host from nmax):
//   W    F[nmax][ld], ld = nmax + 1       J, then W = J - I/(dt*gamma), then its LU
//                                         (FP64 variant only)
//   vec  double[NV][nmax]                 u, uprev, fsal, du, dT, ks[8] (+ dense[3]);
//   pd   double[32]                       parameters (FP64)
//   fsc  float[nmax + 32]                 FP32 copies of uprev and p (mixed variant)
//   piv  int[nmax], perm int[nmax]        pivots and the equivalent row gather
//   gjbuf float[51 NP + 16]               Gauss-Jordan panels, pivot rows, scales and
//                                         mat-vec result (mixed variant; NP = 16 R,
//                                         R = ceil(nmax/16); layout in odemi_gj.h)
//   red  double[32] + int[32]             reductions / broadcasts
// FP64 variant: W is FP64 in LDS (132 KB at n = 128, one workgroup per CU),
// factored by LU with partial pivoting.
// Mixed variant (n <= 128): the FP32 Jacobian is evaluated into a per-workgroup
// global scratch block, W = J - I/(dt*gamma) is held in the workgroup's
// registers (R = ceil(n/16) per trajectory) and inverted in place by swap-free
// Gauss-Jordan with partial pivoting (odemi_gj.h: panels of 16 columns on one
// wave, rank-16 updates on the matrix cores), so each stage solve is one
// parallel matrix-vector product instead of two serial substitutions.  Every stage solve
// gets `refine_sweeps` FP64 iterative-refinement sweeps whose residual
// b - (J x - x/(dt*gamma)) uses the traced Jacobian-vector product in FP64, so
// the refined solution is that of the FP64 system.  LDS use is a few tens of KB,
// so several trajectories share a CU.
#pragma once
#ifdef ODEMI_WITH_WG

#ifndef ODEMI_WG_THREADS
#define ODEMI_WG_THREADS 256
#endif
#define ODEMI_WG_ROWS ((ODEMI_N + 63) / 64)

#include "odemi_gj.h"

namespace odemi {

__host__ __device__ inline int wg_nvec(bool saves) { return 13 + (saves ? 3 : 0); }
constexpr int kWGMaxNP = 32;

struct WGPtrs {
    void* W;
    double *u, *uprev, *fsal, *du, *dT, *lt, *ks, *dense;
    double* pd;
    float* fsc;
    float* pf;
    int* piv;
    int* perm;   // row gather: after the LU, (P b)[i] = b[perm[i]]; Gauss-Jordan: pivot row of step k
    float* gjbuf;   // Gauss-Jordan buffers; zbuf = gjbuf + gj_zbuf_off<R>()
    double* red;
    int* ired;
    int ld;
};

// Gauss-Jordan LDS buffer (floats), NP = 16 R, R = ceil(nmax / 16): 51 NP +
// 16 for the blocked form (two panels [NP][16], pivot rows [16][NP], scales,
// pivot steps, mat-vec result, the current pivot row), which covers the
// element-wise form's 6 NP + 2 R NP (column k and k+1 [2][NP], row staging
// [2][R][NP], scales, result).
__host__ __device__ inline int wg_gj_floats(int nmax, bool gj) {
    const int r = (nmax + 15) / 16, np16 = 16 * r;
    return gj ? 51 * np16 + 16 : 0;
}

// w_in_lds: W lives in LDS (LU variants); the Gauss-Jordan variant keeps it in registers.
template <typename F>
__device__ __forceinline__ WGPtrs wg_layout(unsigned char* base, int nmax, bool saves, bool w_in_lds) {
    WGPtrs L;
    L.ld = nmax + 1;
    size_t off = 0;
    L.W = base;
    if (w_in_lds) off += ((size_t)nmax * L.ld * sizeof(F) + 15) & ~(size_t)15;
    double* v = (double*)(base + off);
    L.u = v; L.uprev = v + nmax; L.fsal = v + 2 * nmax; L.du = v + 3 * nmax; L.dT = v + 4 * nmax;
    L.ks = v + 5 * nmax; L.lt = L.ks; L.dense = v + 13 * nmax;
    off += (size_t)wg_nvec(saves) * nmax * 8;
    L.pd = (double*)(base + off);
    off += (size_t)kWGMaxNP * 8;
    L.fsc = (float*)(base + off);
    L.pf = L.fsc + nmax;
    off += (size_t)(nmax + 32) * 4;
    L.piv = (int*)(base + off);
    off += (size_t)nmax * 4;
    L.perm = (int*)(base + off);
    off += (size_t)nmax * 4;
    off = (off + 15) & ~(size_t)15;
    L.gjbuf = (float*)(base + off);
    off += (size_t)wg_gj_floats(nmax, !w_in_lds) * 4;
    off = (off + 15) & ~(size_t)15;
    L.red = (double*)(base + off);
    L.ired = (int*)(L.red + 32);
    return L;
}

// Block-wide argmax of v over threads (larger value wins, ties -> smaller index).
__device__ __forceinline__ int wg_argmax(double v, int idx, double* red, int* ired) {
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(v, off);
        int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if (lane == 0) { red[w] = v; ired[w] = idx; }
    __syncthreads();
    double bv = red[0];
    int bi = ired[0];
    const int nw = blockDim.x >> 6;
    for (int q = 1; q < nw; q++) {
        double ov = red[q];
        int oi = ired[q];
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    return bi;
}

// Broadcast a value held by lane `src` (wave-uniform) to the whole wave.
__device__ __forceinline__ float wave_bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double wave_bcast(double v, int src) {
    i64 b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((i64)hi << 32) | (i64)(unsigned)lo);
}

// In-place LU with partial pivoting of the n x n leading block of A (stride ld).
#define ODEMI_WG_CB ((ODEMI_N + 15) / 16)
template <typename F>
__device__ void wg_lu(F* A, int ld, int n, int* piv, int* perm, double* red, int* ired) {
    const int tid = threadIdx.x, NT = blockDim.x;
    for (int i = tid; i < n; i += NT) perm[i] = i;
    for (int k = 0; k < n; k++) {
        // pivot: first maximum of |A[i][k]|, i >= k (a NaN diagonal keeps p = k)
        double v = -1.0;
        int idx = 0x7fffffff;
        for (int i = k + tid; i < n; i += NT) {
            double a = fabs((double)A[i * ld + k]);
            if (!isnan(a) && (a > v || (a == v && i < idx))) { v = a; idx = i; }
        }
        int p = wg_argmax(v, idx, red, ired);
        if (isnan((double)A[k * ld + k]) || p == 0x7fffffff) p = k;
        if (tid == 0) {
            piv[k] = p;
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
        }
        if (p != k) {
            for (int c = tid; c < n; c += NT) {
                F a = A[k * ld + c];
                A[k * ld + c] = A[p * ld + c];
                A[p * ld + c] = a;
            }
        }
        __syncthreads();
        // multipliers and trailing update on a 16 x (NT/16) thread grid.  The 16
        // threads of one row are lanes of one wave: they all read the unscaled
        // A[i][k] before lane tj == 0 stores the multiplier.  Each thread keeps
        // its slice of the pivot row in registers.
        const F inv = (F)1 / A[k * ld + k];
        const int ti = tid >> 4, tj = tid & 15, TI = NT >> 4;
        F pr[ODEMI_WG_CB];
#pragma unroll
        for (int b = 0; b < ODEMI_WG_CB; b++) {
            const int c = k + 1 + tj + 16 * b;
            pr[b] = (c < n) ? A[k * ld + c] : (F)0;
        }
        for (int i = k + 1 + ti; i < n; i += TI) {
            F* row = A + i * ld;
            const F l = row[k] * inv;
            F a[ODEMI_WG_CB];
#pragma unroll
            for (int b = 0; b < ODEMI_WG_CB; b++) {
                const int c = k + 1 + tj + 16 * b;
                a[b] = (c < n) ? row[c] : (F)0;
            }
#pragma unroll
            for (int b = 0; b < ODEMI_WG_CB; b++) {
                const int c = k + 1 + tj + 16 * b;
                if (c < n) row[c] = fma(-l, pr[b], a[b]);
            }
            if (tj == 0) row[k] = l;
        }
        __syncthreads();
    }
}

// x <- A^{-1} x using the LU of wg_lu.  x is FP64 in LDS; the substitutions run
// in precision F on wave 0 (rows held in registers, the solved entry broadcast
// with readlane), in LAPACK's column order.  All threads must call.
template <typename F>
__device__ void wg_lu_solve(const F* A, int ld, int n, const int* perm, double* x) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        // row interchanges of the LU applied as one gather (same result as the
        // sequential swaps b[k] <-> b[piv[k]], k = 0..n-1)
        F b[ODEMI_WG_ROWS];
#pragma unroll
        for (int r = 0; r < ODEMI_WG_ROWS; r++) {
            int i = lane + 64 * r;
            b[r] = (i < n) ? (F)x[perm[i]] : (F)0;
        }
        // forward, unit lower triangle (column order); the solved entry is
        // broadcast with readlane (uniform lane), the factor column is read
        // one step ahead so the dependent chain is readlane + fma
        F Lc[ODEMI_WG_ROWS], Ln[ODEMI_WG_ROWS];
#pragma unroll
        for (int r = 0; r < ODEMI_WG_ROWS; r++) {
            int i = lane + 64 * r;
            Lc[r] = (i < n && n > 1) ? A[i * ld + 0] : (F)0;
        }
        for (int c = 0; c < n - 1; c++) {
#pragma unroll
            for (int r = 0; r < ODEMI_WG_ROWS; r++) {
                int i = lane + 64 * r;
                Ln[r] = (i < n && c + 1 < n - 1) ? A[i * ld + c + 1] : (F)0;
            }
            F bc = b[0];
#pragma unroll
            for (int r = 1; r < ODEMI_WG_ROWS; r++) if ((c >> 6) == r) bc = b[r];
            const F xc = wave_bcast(bc, c & 63);
#pragma unroll
            for (int r = 0; r < ODEMI_WG_ROWS; r++) {
                int i = lane + 64 * r;
                if (i > c && i < n) b[r] = fma(-Lc[r], xc, b[r]);
                Lc[r] = Ln[r];
            }
        }
        // backward, upper triangle (column order)
        F Uc[ODEMI_WG_ROWS], Un[ODEMI_WG_ROWS];
        F dc = A[(n - 1) * ld + (n - 1)], dn = (F)1;
#pragma unroll
        for (int r = 0; r < ODEMI_WG_ROWS; r++) {
            int i = lane + 64 * r;
            Uc[r] = (i < n) ? A[i * ld + (n - 1)] : (F)0;
        }
        for (int c = n - 1; c >= 0; c--) {
            if (c > 0) {
                dn = A[(c - 1) * ld + (c - 1)];
#pragma unroll
                for (int r = 0; r < ODEMI_WG_ROWS; r++) {
                    int i = lane + 64 * r;
                    Un[r] = (i < n) ? A[i * ld + (c - 1)] : (F)0;
                }
            }
            F bc = b[0];
#pragma unroll
            for (int r = 1; r < ODEMI_WG_ROWS; r++) if ((c >> 6) == r) bc = b[r];
            const F xc = wave_bcast(bc, c & 63) / dc;
#pragma unroll
            for (int r = 0; r < ODEMI_WG_ROWS; r++) {
                int i = lane + 64 * r;
                if (i == c) b[r] = xc;
                else if (i < c) b[r] = fma(-Uc[r], xc, b[r]);
                Uc[r] = Un[r];
            }
            dc = dn;
        }
#pragma unroll
        for (int r = 0; r < ODEMI_WG_ROWS; r++) {
            int i = lane + 64 * r;
            if (i < n) x[i] = (double)b[r];
        }
    }
    __syncthreads();
}

// Stage solve W x = b (b in `x` on entry; x must be visible to all threads).
// FP64: one LU solve.  Mixed: FP32 LU solve plus refinement sweeps with the
// FP64 residual r = b - J x + sigma x (J x from the traced Jacobian-vector
// product at uprev).  Scratch: rb (copy of b), r, d -- three free vectors.
template <typename F>
__device__ void wg_wsolve(const WGPtrs& S, int n, double* x, double* rb, double* r, double* d,
                          const double* uprev, const double* p, double t, double sigma, int sweeps) {
    const F* A = (const F*)S.W;
    if constexpr (sizeof(F) == sizeof(double)) {
        wg_lu_solve<F>(A, S.ld, n, S.perm, x);
    } else {
        const int tid = threadIdx.x, NT = blockDim.x;
        for (int i = tid; i < n; i += NT) rb[i] = x[i];
        __syncthreads();
        wg_lu_solve<F>(A, S.ld, n, S.perm, x);
        for (int sw = 0; sw < sweeps; sw++) {
            odemi_jvp_wg<double>(r, x, uprev, p, t, n, tid, NT);
            __syncthreads();
            for (int i = tid; i < n; i += NT) d[i] = fma(sigma, x[i], rb[i] - r[i]);
            __syncthreads();
            wg_lu_solve<F>(A, S.ld, n, S.perm, d);
            for (int i = tid; i < n; i += NT) x[i] = x[i] + d[i];
            __syncthreads();
        }
    }
}

#if ODEMI_GJ
// Stage solve W x = b for the Gauss-Jordan variant (b in x on entry), with the
// same FP64 refinement sweeps as wg_wsolve.  gjbuf holds gj_invert<R>'s
// buffers (sized for the largest R), zbuf at gj_zbuf_off<R>().
template <int R>
__device__ __forceinline__ void wg_wsolve_gj(const GJTile<R>& G, const WGPtrs& S, int n, double* x, double* rb, double* r, double* d,
                             const double* uprev, const double* p, double t, double sigma, int sweeps) {
    const int tid = threadIdx.x, NT = blockDim.x;
    float* zbuf = S.gjbuf + gj_zbuf_off<R>();
    for (int i = tid; i < n; i += NT) rb[i] = x[i];
    __syncthreads();
    gj_solve<R>(G, n, S.perm, x, zbuf);
    for (int sw = 0; sw < sweeps; sw++) {
        odemi_jvp_wg<double>(r, x, uprev, p, t, n, tid, NT);
        __syncthreads();
        for (int i = tid; i < n; i += NT) d[i] = fma(sigma, x[i], rb[i] - r[i]);
        __syncthreads();
        gj_solve<R>(G, n, S.perm, d, zbuf);
        for (int i = tid; i < n; i += NT) x[i] = x[i] + d[i];
        __syncthreads();
    }
}
#endif  // ODEMI_GJ

__device__ __forceinline__ void wg_save_point(Ctl& c, double& tlast, const Params& P, i64 j, double t,
                                              const double* v, int n) {
    const int s = c.nsaved;
    if (s < P.max_saves) {
        if (threadIdx.x == 0) P.t_out[(i64)s * P.N + j] = t;
        for (int i = threadIdx.x; i < n; i += blockDim.x) P.u_out[((i64)s * P.nmax + i) * P.N + j] = v[i];
    }
    c.nsaved = s + 1;
    tlast = t;
}

// Index-order fma-chain norms, evaluated by one thread (bit-identical to the
// sequential restatement).
__device__ __forceinline__ double wg_norm_div(const double* a, const double* sk, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) { double x = a[i] / sk[i] * 1.0; s = fma(x, x, s); }
    return sqrt(s / (double)(n > 1 ? n : 1));
}
__device__ __forceinline__ double wg_norm_diff_div(const double* a, const double* b, const double* sk, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) { double x = (a[i] - b[i]) / sk[i] * 1.0; s = fma(x, x, s); }
    return sqrt(s / (double)(n > 1 ? n : 1));
}

// initdt.jl:17-298, cooperative: RHS evaluations by the whole workgroup, the
// scalar logic and norms on thread 0, result broadcast.  Uses ks[0], du, dT,
// fsal as scratch (sk, f0, f1, u1).
__device__ double wg_initdt(const WGPtrs& S, const Ctl& c, const double* u0, const double* p, int n,
                            double reltol, double abstol, double dtmin_opt) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const double t = c.t, tdir = c.tdir;
    const double dtmax_tdir = jl_min(fabs(c.dtmax), fabs(c.tnext - t));   // first tstop
    const double dtmin = nextfloat_pos(jl_max(dtmin_opt, jl_eps(t)));
    const double smalldt = jl_max(dtmin, 1e-6);
    double* sk = S.lt;
    double* f0 = S.du;
    double* f1 = S.dT;
    double* u1 = S.fsal;
    double* red = S.red;
    for (int i = tid; i < n; i += NT) sk[i] = fma(fabs(u0[i]), reltol, abstol);
    odemi_f_wg<double>(f0, u0, p, t, n, tid, NT);
    __syncthreads();
    if (tid == 0) {
        double d0 = wg_norm_div(u0, sk, n);
        double d1 = wg_norm_div(f0, sk, n);
        double res = 0.0, need = 0.0, dt0 = 0.0;
        if (isnan(d1)) {
            res = tdir * dtmin;
        } else {
            dt0 = ((d0 < 1e-5) | (d1 < 1e-5)) ? smalldt : (d0 / d1) / 100.0;
            dt0 = jl_min(dt0, dtmax_tdir);
            if (dt0 < 10.0 * 2.220446049250313e-16) res = tdir * jl_max(smalldt, dtmin);
            else need = 1.0;
        }
        red[20] = res; red[21] = need; red[22] = dt0; red[23] = d1;
    }
    __syncthreads();
    double res = red[20];
    const bool need = red[21] != 0.0;
    const double dt0 = red[22];
    __syncthreads();
    if (!need) return res;
    const double dt0_tdir = tdir * dt0;
    for (int i = tid; i < n; i += NT) u1[i] = fma(dt0_tdir, f0[i], u0[i]);
    __syncthreads();
    odemi_f_wg<double>(f1, u1, p, t + dt0_tdir, n, tid, NT);
    __syncthreads();
    if (tid == 0) {
        bool same = true;
        for (int i = 0; i < n; i++) same = same && (f0[i] == f1[i]);
        double r;
        if (same) {
            r = tdir * jl_max(dtmin, 100.0 * dt0);
        } else {
            double d1 = red[23];
            double d2 = wg_norm_diff_div(f1, f0, sk, n) / dt0 * 1.0;
            double mx = jl_max(d1, d2);
            double dt1;
            if (mx < 1e-15) dt1 = jl_max(1e-6, dt0 * 0.001);
            else dt1 = det_exp(-(2.0 + det_log(mx) / kLN10) / 5.0 * kLN10);
            r = tdir * jl_max(dtmin, jl_min(jl_min(100.0 * dt0, dt1), dtmax_tdir));
        }
        red[20] = r;
    }
    __syncthreads();
    res = red[20];
    __syncthreads();
    return res;
}

// ---------------------------------------------------------------------------
// Workgroup Rodas5P, in three phases so the batch kernel and the serving
// kernels share every line of the step: wg_begin (u0, parameters, controller,
// initial dt), wg_run (step attempts), wg_finish (postamble and write-out).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wg_load_params(const WGPtrs& S, const Params& P, i64 j) {
    for (int k = threadIdx.x; k < ODEMI_NP; k += blockDim.x) {
        const double v = (k < P.np) ? P.p[(i64)k * P.N + j] : 0.0;
        S.pd[k] = v;
        S.pf[k] = (float)v;
    }
}

template <bool GJ>
__device__ __forceinline__ void wg_begin(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double& tlast,
                                         float* Jg) {
    const int tid = threadIdx.x, NT = blockDim.x;
    wg_load_params(S, P, j);
    for (int i = tid; i < n; i += NT) { double v = P.u0[(i64)i * P.N + j]; S.u[i] = v; S.uprev[i] = v; S.dT[i] = 0.0; }
    ctl_init(c, P, j);
    tlast = 0.0;
#if ODEMI_GJ
    // the traced Jacobian writes the same entries on every call for a given n,
    // so the rest of the block is cleared once per trajectory
    if constexpr (GJ)
        for (int e = tid; e < n * ODEMI_GJ_LD; e += NT) Jg[e] = 0.0f;
#endif
    __syncthreads();
    if (P.save_start) wg_save_point(c, tlast, P, j, c.t, S.u, n);
    if (P.dt_user != 0.0) {
        c.dt = user_dt(c, P);
    } else {
        c.dt = wg_initdt(S, c, S.u, S.pd, n, P.reltol, P.abstol, P.dtmin);
        c.nf += 2;
    }
    c.dtpropose = c.dt;
}

#if defined(ODEMI_NDCB_WG) || defined(ODEMI_NCCB_WG)
#define ODEMI_WG_CALLBACKS 1
#endif
#ifdef ODEMI_NCCB_WG
#define ODEMI_WG_EVENTS 1
#else
#define ODEMI_WG_EVENTS 0
#endif

#ifdef ODEMI_WG_CALLBACKS
// ---------------------------------------------------------------------------
// Callbacks in the workgroup kernels (handle_callbacks!, integrator_utils.jl:
// 1081-1132; lib/DiffEqBase/src/callbacks.jl:157-683), the same sequence as
// the lane engines: continuous callbacks first (the earliest event wins, the
// first callback on ties), then the discrete ones in order, then savevalues!
// unless a callback already ran it.  Conditions and affects are traced per
// state size (OdemiCcbWg<K>, OdemiDcbWg<K>) and run by thread 0 on the state in
// LDS; every thread holds the same controller state, so the branches on their
// results are uniform.
// ---------------------------------------------------------------------------
// Julia's sign(): -1, 0, 1, NaN for NaN
__device__ __forceinline__ double wg_sign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : (x == 0 ? 0.0 : x)); }

// The Rosenbrock dense output of the accepted step at theta into v (all
// threads; rosenbrock_interpolants.jl:172-206, the form of the saveat path).
__device__ __forceinline__ void wg_interp(const WGPtrs& S, int n, int nmax, double th, double* v) {
    const double th1 = 1.0 - th;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double in1 = fma(th, S.dense[2 * nmax + i], S.dense[nmax + i]);
        const double in2 = fma(th, in1, S.dense[i]);
        const double in3 = fma(th1, in2, S.u[i]);
        v[i] = fma(th, in3, th1 * S.uprev[i]);
    }
    __syncthreads();
}

#ifdef ODEMI_NCCB_WG
// get_condition (callbacks.jl:91-146): u at t, uprev at tprev, else the dense
// output (into S.du, free after the step); every call counts in ncondition.
template <int CB>
__device__ double wg_ccb_condition(const WGPtrs& S, Ctl& c, int n, int nmax, double abst) {
    c.ncond++;
    const double* x = S.u;
    if (abst != c.t) {
        if (abst == c.tprev) {
            x = S.uprev;
        } else {
            wg_interp(S, n, nmax, (abst - c.tprev) / c.dt, S.du);
            x = S.du;
        }
    }
    if (threadIdx.x == 0) S.red[27] = (double)OdemiCcbWg<CB>::template cond<double>(x, S.pd, abst, n);
    __syncthreads();
    const double r = S.red[27];
    __syncthreads();
    return r;
}

template <int CB>
__device__ __forceinline__ bool wg_ccb_is_event(double prev_sign, double next_sign) {
    return ((prev_sign < 0 && OdemiCcbWg<CB>::has_affect) || (prev_sign > 0 && OdemiCcbWg<CB>::has_affect_neg)) &&
           prev_sign * next_sign <= 0;
}

// find_callback_time (callbacks.jl:364-412) with nudge_tprev (414-426) and the
// interp_points checks; the root bracketed down to adjacent floats by
// bisection (as the lane engines and the restatement do).
template <int CB>
__device__ bool wg_ccb_find_time(const WGPtrs& S, Ctl& c, int n, int nmax, double& cb_t, double& bottom_sign,
                                 double& residual) {
    typedef OdemiCcbWg<CB> C;
    double bottom_t = c.tprev;
    double bc = wg_ccb_condition<CB>(S, c, n, nmax, bottom_t);
    if (c.event_last == CB + 1) {
        if (fabs(bc - c.last_event_error) <= C::abstol) bottom_t = c.tprev + c.dt * C::repeat_nudge;
        bc = wg_ccb_condition<CB>(S, c, n, nmax, bottom_t);
    }
    const double bs = wg_sign(bc);
    double top_t = c.t;
    double ts = wg_sign(wg_ccb_condition<CB>(S, c, n, nmax, top_t));
    bool ev = wg_ccb_is_event<CB>(bs, ts);
    if (C::interp_points != 0 && !ev) {
        const double h = (c.t - c.tprev) / (double)(C::interp_points - 1);
        for (int i = 2; i <= C::interp_points; i++) {
            top_t = (i == C::interp_points) ? c.t : fma((double)(i - 1), h, c.tprev);
            ts = wg_sign(wg_ccb_condition<CB>(S, c, n, nmax, top_t));
            ev = wg_ccb_is_event<CB>(bs, ts);
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
        const double sm = wg_sign(wg_ccb_condition<CB>(S, c, n, nmax, m));
        if (sm == 0.0) { l = m; r = m; break; }
        if (sm == bs) l = m;
        else r = m;
    }
    cb_t = (C::rootfind == 1) ? r : l;
    residual = wg_ccb_condition<CB>(S, c, n, nmax, cb_t);
    return true;
}
#ifndef ODEMI_VCB_MAXLEN_WG
#define ODEMI_VCB_MAXLEN_WG 1
#endif
// get_condition for a VectorContinuousCallback: its len values at abst, by every
// thread from the state in LDS (uniform values); one condition call each time.
template <int CB>
__device__ void wg_vcb_condition(const WGPtrs& S, Ctl& c, int n, int nmax, double abst, double* out) {
    c.ncond++;
    const double* x = S.u;
    if (abst != c.t) {
        if (abst == c.tprev) {
            x = S.uprev;
        } else {
            wg_interp(S, n, nmax, (abst - c.tprev) / c.dt, S.du);
            x = S.du;
        }
    }
    OdemiCcbWg<CB>::template vcond<double>(out, x, S.pd, abst, n);
    __syncthreads();   // S.du is rewritten by the next interpolation
}
// findall_events! (callbacks.jl:499-521): event i when the bottom sign is
// nonzero and bottom sign * value <= 0
template <int V>
__device__ __forceinline__ unsigned wg_vcb_events(const double* bs, const double* val) {
    unsigned m = 0u;
#pragma unroll
    for (int i = 0; i < V; i++)
        if (bs[i] != 0 && bs[i] * val[i] <= 0) m |= 1u << i;
    return m;
}
// find_callback_time for a VectorContinuousCallback (callbacks.jl:239-362), as
// the lane engines do it: the components that fired last time re-read at the
// nudged tprev (nudge_tprev on the value closest to zero among them), the
// interp_points checks unless every component has an event at t, each event
// component rooted to adjacent floats, the earliest root winning and the
// components rooted at that same time simultaneous (direction -sign(bottom)).
template <int CB>
__device__ bool wg_vcb_find_time(const WGPtrs& S, Ctl& c, int n, int nmax, double& cb_t, double* ev, double& residual) {
    typedef OdemiCcbWg<CB> C;
    constexpr int V = C::vlen;
    double bc[V], bs[V], top[V], tmp[V];
    const double bottom_t = c.tprev;
    wg_vcb_condition<CB>(S, c, n, nmax, bottom_t, bc);
#pragma unroll
    for (int i = 0; i < V; i++) { bs[i] = wg_sign(bc[i]); ev[i] = 0.0; }
    double nudged_t = bottom_t;
    const unsigned prev = (c.event_last == CB + 1) ? c.sim_mask : 0u;
    if (c.event_last == CB + 1) {
        double minval = 0.0, minabs = __builtin_inf();
#pragma unroll
        for (int i = 0; i < V; i++)
            if (((prev >> i) & 1u) && fabs(bc[i]) < minabs) { minabs = fabs(bc[i]); minval = bc[i]; }
        if (fabs(minval - c.last_event_error) <= C::abstol) nudged_t = c.tprev + c.dt * C::repeat_nudge;
        wg_vcb_condition<CB>(S, c, n, nmax, nudged_t, tmp);
#pragma unroll
        for (int i = 0; i < V; i++)
            if ((prev >> i) & 1u) bs[i] = wg_sign(tmp[i]);
    }
    double top_t = c.t;
    wg_vcb_condition<CB>(S, c, n, nmax, top_t, top);
    unsigned mask = wg_vcb_events<V>(bs, top);
    constexpr unsigned all = (V >= 32) ? 0xffffffffu : ((1u << V) - 1u);
    if (C::interp_points != 0 && mask != all) {
        const double h = (c.t - c.tprev) / (double)(C::interp_points - 1);
        for (int i = 2; i <= C::interp_points; i++) {
            top_t = (i == C::interp_points) ? c.t : fma((double)(i - 1), h, c.tprev);
            wg_vcb_condition<CB>(S, c, n, nmax, top_t, top);
            mask = wg_vcb_events<V>(bs, top);
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
    double cbt = (c.tdir > 0) ? __longlong_as_double(__double_as_longlong(top_t) + (top_t >= 0.0 ? 1 : -1))
                              : __longlong_as_double(__double_as_longlong(top_t) + (top_t > 0.0 ? -1 : 1)), res = 0.0;
    if (top_t == 0.0) cbt = (c.tdir > 0) ? __longlong_as_double(1) : -__longlong_as_double(1);
#pragma unroll
    for (int idx = 0; idx < V; idx++) {
        if (!((mask >> idx) & 1u)) continue;
        double cbi;
        if (wg_sign(top[idx]) == 0.0) {
            cbi = top_t;
        } else {
            double l = ((prev >> idx) & 1u) ? nudged_t : bottom_t, r = top_t;
            for (int it = 0; it < 2200; it++) {
                const double m = l + 0.5 * (r - l);
                if (m == l || m == r) break;
                wg_vcb_condition<CB>(S, c, n, nmax, m, tmp);
                const double sm = wg_sign(tmp[idx]);
                if (sm == 0.0) { l = m; r = m; break; }
                if (sm == bs[idx]) l = m;
                else r = m;
            }
            cbi = (C::rootfind == 1) ? r : l;
        }
        if (c.tdir * cbi < c.tdir * cbt) {
#pragma unroll
            for (int i = 0; i < V; i++) ev[i] = 0.0;
        }
        if (c.tdir * cbi <= c.tdir * cbt) {
            cbt = cbi;
            wg_vcb_condition<CB>(S, c, n, nmax, cbi, tmp);
            res = tmp[idx];
            ev[idx] = -bs[idx];
        }
    }
    cb_t = cbt; residual = res;
    return true;
}

// _ode_addsteps! for RosenbrockCache (stiff_addsteps.jl:140-224) after
// change_t_via_interpolation!, in the lane engines' form: the stages of the
// step from uprev at tprev with the new dt, the step's dT and fsalfirst
// f(uprev, tprev) (re-evaluated), W = J - I/(dt gamma) from J at (uprev,
// tprev) (the step's Jacobian: FP64 re-evaluated into LDS, or the FP32 block
// of this attempt for Gauss-Jordan) refactored; no @muladd there, so products
// and sums round separately except linsolve_tmp = fsalfirst + dtgamma dT;
// dtC = C ./ dt; the dense rows from the new stages.  The dense rows serve as
// solve scratch until they are rebuilt.  No stats change.
template <typename F, bool GJ, int R, typename GT>
__device__ void wg_addsteps(const WGPtrs& S, const Params& P, const Ctl& c, int n, GT& G, float* Jg) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const int nmax = P.nmax, ld = S.ld;
    const double t = c.tprev, dt = c.dt;
    const double dtgamma = dt * RodasTab::gamma;
    const double sigma = 1.0 / dtgamma;
    double* ks = S.ks;
    double* sc0 = S.dense;
    double* sc1 = S.dense + nmax;
    double* sc2 = S.dense + 2 * nmax;
    (void)G; (void)Jg; (void)ld;
    odemi_f_wg<double>(S.fsal, S.uprev, S.pd, t, n, tid, NT);
    __syncthreads();
    for (int i = tid; i < n; i += NT) ks[i] = fma(dtgamma, S.dT[i], S.fsal[i]);
    if constexpr (GJ) {
#if ODEMI_GJ
        gj_load<R>(G, Jg, n, sigma);
        gj_invert<R>(G, n, S.perm, S.gjbuf);
#endif
    } else {
        F* W = (F*)S.W;
        for (int e = tid; e < n * ld; e += NT) W[e] = (F)0;
        if (sizeof(F) != sizeof(double))
            for (int i = tid; i < n; i += NT) S.fsc[i] = (float)S.uprev[i];
        __syncthreads();
        if constexpr (sizeof(F) == sizeof(double)) odemi_jac_wg<double>((double*)W, ld, S.uprev, S.pd, t, n, tid, NT);
        else odemi_jac_wg<float>((float*)W, ld, S.fsc, S.pf, (float)t, n, tid, NT);
        __syncthreads();
        for (int i = tid; i < n; i += NT) W[i * ld + i] = (F)fma(-1.0, sigma, (double)W[i * ld + i]);
        __syncthreads();
        wg_lu<F>(W, ld, n, S.piv, S.perm, S.red, S.ired);
    }
    __syncthreads();
    auto solve = [&](double* x) {
        if constexpr (GJ) {
#if ODEMI_GJ
            wg_wsolve_gj<R>(G, S, n, x, sc0, sc1, sc2, S.uprev, S.pd, t, sigma, P.refine_sweeps);
#endif
        } else {
            wg_wsolve<F>(S, n, x, sc0, sc1, sc2, S.uprev, S.pd, t, sigma, P.refine_sweeps);
        }
        for (int i = tid; i < n; i += NT) x[i] = -x[i];
        __syncthreads();
    };
    solve(ks);
    for (int s = 1; s < 8; s++) {
        for (int i = tid; i < n; i += NT) {
            double acc = S.uprev[i];
            for (int q = 0; q < s; q++) acc = acc + kRA[s][q] * ks[q * nmax + i];
            S.fsal[i] = acc;
        }
        __syncthreads();
        odemi_f_wg<double>(S.du, S.fsal, S.pd, t + kRc[s] * dt, n, tid, NT);
        __syncthreads();
        const double dtds = dt * kRd[s];
        double* x = ks + s * nmax;
        for (int i = tid; i < n; i += NT) {
            double acc = S.du[i] + dtds * S.dT[i];
            for (int q = 0; q < s; q++) acc = acc + (kRC[s][q] / dt) * ks[q * nmax + i];
            x[i] = acc;
        }
        __syncthreads();
        solve(x);
    }
    for (int r = 0; r < 3; r++)
        for (int i = tid; i < n; i += NT) {
            double acc = 0.0;
            for (int q = 0; q < 8; q++) acc = acc + kRH[r][q] * ks[q * nmax + i];
            S.dense[r * nmax + i] = acc;
        }
    __syncthreads();
}

// apply_callback! (callbacks.jl:557-640): set_proposed_dt!, the state moved to
// the event time through the dense output with the stages recomputed, saves,
// affect! on an upcrossing / affect_neg! on a downcrossing.  Returns terminate!.
template <int CB, typename F, bool GJ, int R, typename GT, typename SaveFn>
__device__ bool wg_ccb_apply(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double& tlast, double cb_t,
                             double bs, const double* ev, GT& G, float* Jg, SaveFn& savevalues, bool& saved_in_cb) {
    typedef OdemiCcbWg<CB> C;
    const int tid = threadIdx.x, nmax = P.nmax;
    c.dtpropose = c.tdir * jl_max(nextfloat_pos(P.dtmin), c.tdir * c.dt);
    if (cb_t != c.t) {
        wg_interp(S, n, nmax, (cb_t - c.tprev) / c.dt, S.du);
        for (int i = tid; i < n; i += blockDim.x) S.u[i] = S.du[i];
        __syncthreads();
        c.t = cb_t;
        c.dt = c.t - c.tprev;
        wg_addsteps<F, GJ, R>(S, P, c, n, G, Jg);
        c.reeval_fsal = 1;
    }
    const bool savedexactly = savevalues();
    saved_in_cb = true;
    if (C::save0 && !savedexactly) wg_save_point(c, tlast, P, j, c.t, S.u, n);
    __syncthreads();
    bool modified = true, term = false;
    if constexpr (C::vlen > 0) {
        // VectorContinuousCallback: affect!(integrator, simultaneous_events)
        if constexpr (C::has_affect != 0) {
            if (tid == 0) S.red[28] = C::template vaffect<double>(S.u, S.pd, c.t, ev, n) ? 1.0 : 0.0;
            __syncthreads();
            term = S.red[28] != 0.0;
        } else {
            modified = false;
        }
    } else if (bs < 0) {
        if constexpr (C::has_affect != 0) {
            if (tid == 0) C::template affect<double>(S.u, S.pd, c.t, n);
            term = C::term != 0;
        } else {
            modified = false;
        }
    } else if (bs > 0) {
        if constexpr (C::has_affect_neg != 0) {
            if (tid == 0) C::template affect_neg<double>(S.u, S.pd, c.t, n);
            term = C::term_neg != 0;
        } else {
            modified = false;
        }
    }
    __syncthreads();
    if (modified) {
        c.reeval_fsal = 1;
        if (C::save1) { wg_save_point(c, tlast, P, j, c.t, S.u, n); saved_in_cb = true; }
    }
    return term;
}

template <int CB>
__device__ __forceinline__ void wg_ccb_find_first(const WGPtrs& S, Ctl& c, int n, int nmax, int& win, double& tmin,
                                                  double& bs_win, double& res_win, double* ev_win) {
    if constexpr (CB < ODEMI_NCCB_WG) {
        typedef OdemiCcbWg<CB> C;
        double cbt, bs = 0.0, rs;
        if constexpr (C::vlen > 0) {
            double ev[C::vlen];
            if (wg_vcb_find_time<CB>(S, c, n, nmax, cbt, ev, rs) && (win < 0 || c.tdir * cbt < c.tdir * tmin)) {
                win = CB; tmin = cbt; bs_win = bs; res_win = rs;
#pragma unroll
                for (int i = 0; i < ODEMI_VCB_MAXLEN_WG; i++) ev_win[i] = (i < C::vlen) ? ev[i < C::vlen ? i : 0] : 0.0;
            }
        } else {
            if (wg_ccb_find_time<CB>(S, c, n, nmax, cbt, bs, rs) && (win < 0 || c.tdir * cbt < c.tdir * tmin)) {
                win = CB; tmin = cbt; bs_win = bs; res_win = rs;
#pragma unroll
                for (int i = 0; i < ODEMI_VCB_MAXLEN_WG; i++) ev_win[i] = 0.0;
            }
        }
        wg_ccb_find_first<CB + 1>(S, c, n, nmax, win, tmin, bs_win, res_win, ev_win);
    }
}
template <int CB, typename F, bool GJ, int R, typename GT, typename SaveFn>
__device__ __forceinline__ bool wg_ccb_apply_winner(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c,
                                                    double& tlast, int win, double cbt, double bs, const double* ev,
                                                    GT& G, float* Jg, SaveFn& savevalues, bool& saved_in_cb) {
    if constexpr (CB < ODEMI_NCCB_WG) {
        if (win == CB)
            return wg_ccb_apply<CB, F, GJ, R>(S, P, j, n, c, tlast, cbt, bs, ev, G, Jg, savevalues, saved_in_cb);
        return wg_ccb_apply_winner<CB + 1, F, GJ, R>(S, P, j, n, c, tlast, win, cbt, bs, ev, G, Jg, savevalues,
                                                     saved_in_cb);
    } else {
        return false;
    }
}
#endif  // ODEMI_NCCB_WG

#ifdef ODEMI_NDCB_WG
// apply_discrete_callback! (callbacks.jl:645-683) for each DiscreteCallback in
// order, as the lane engines do.
template <int K, typename SaveFn>
__device__ __forceinline__ void wg_dcb_apply(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double& tlast,
                                             SaveFn& savevalues, bool& saved_in_cb, bool& term) {
    if constexpr (K < ODEMI_NDCB_WG) {
        typedef OdemiDcbWg<K> D;
        c.ncond++;
        if (threadIdx.x == 0) S.red[26] = D::template cond<double>(S.u, S.pd, c.t, n) ? 1.0 : 0.0;
        __syncthreads();
        const bool fire = S.red[26] != 0.0;
        __syncthreads();
        if (fire) {
            const bool savedexactly = savevalues();
            saved_in_cb = true;
            if (D::save0 && !savedexactly) wg_save_point(c, tlast, P, j, c.t, S.u, n);
            __syncthreads();
            if (threadIdx.x == 0) D::template affect<double>(S.u, S.pd, c.t, n);
            __syncthreads();
            c.reeval_fsal = 1;
            if (D::term) term = true;
            if (D::save1) wg_save_point(c, tlast, P, j, c.t, S.u, n);
        }
        wg_dcb_apply<K + 1>(S, P, j, n, c, tlast, savevalues, saved_in_cb, term);
    }
}
#endif

// handle_callbacks!: returns terminate!.
template <typename F, bool GJ, int R, typename GT, typename SaveFn>
__device__ __forceinline__ bool wg_handle_callbacks(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c,
                                                    double& tlast, GT& G, float* Jg, SaveFn& savevalues) {
    bool saved_in_cb = false, term = false;
    (void)G; (void)Jg;
#ifdef ODEMI_NCCB_WG
    {
        int win = -1;
        double tmin = 0.0, bs = 0.0, rs = 0.0;
        double ev[ODEMI_VCB_MAXLEN_WG];
#pragma unroll
        for (int i = 0; i < ODEMI_VCB_MAXLEN_WG; i++) ev[i] = 0.0;
        wg_ccb_find_first<0>(S, c, n, P.nmax, win, tmin, bs, rs, ev);
        if (win < 0) {
            c.event_last = 0;
        } else {
            c.last_event_error = rs;
            c.event_last = win + 1;
            unsigned m = 0u;
#pragma unroll
            for (int i = 0; i < ODEMI_VCB_MAXLEN_WG; i++)
                if (ev[i] != 0.0) m |= 1u << i;
            c.sim_mask = m;
            term = wg_ccb_apply_winner<0, F, GJ, R>(S, P, j, n, c, tlast, win, tmin, bs, ev, G, Jg, savevalues,
                                                    saved_in_cb);
        }
    }
#endif
#ifdef ODEMI_NDCB_WG
    wg_dcb_apply<0>(S, P, j, n, c, tlast, savevalues, saved_in_cb, term);
#endif
    if (!saved_in_cb) savevalues();
    return term;
}
#endif  // ODEMI_WG_CALLBACKS

// Step attempts until the trajectory ends (returns its ReturnCode) or `limit`
// attempts have been made (returns 0; the state is in c and S.u/uprev/dT).
// R: Gauss-Jordan tile (GJ variant only; 16 R >= n).
template <typename F, bool SAVES, bool GJ, int R = 1>
__device__ __forceinline__ int wg_run(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double& tlast,
                                      float* Jg, int limit) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const CtlConst kc = ctl_const(true);
    double* u = S.u;
    double* uprev = S.uprev;
    double* fsal = S.fsal;
    double* du = S.du;
    double* dT = S.dT;
    double* ks = S.ks;
    double* dense = S.dense;
    double* red = S.red;
    const double* pp = S.pd;
    const float* pf = S.pf;
    const int nmax = P.nmax, ld = S.ld;
    F* W = (F*)S.W;
    (void)Jg; (void)W; (void)ld; (void)pf; (void)dense;
#if ODEMI_GJ
    GJTile<R> G;
#endif
    int rc = RC_SUCCESS;
    for (int att = 0; c.tdir * c.t < c.tdir * c.tf; att++) {
        if (att == limit) return 0;   // uniform: every thread holds the same controller state
        // ---- loopheader! ----
        if (c.iter > 0) {
            if (c.accept_step) {
                c.success_iter++;
                for (int i = tid; i < n; i += NT) uprev[i] = u[i];
                c.dt = c.dtpropose;
                apply_step_tstop_clip(c);
            } else {
#ifdef ODEMI_HAS_ISOUT_WG
                if (c.isout) c.dt = c.dt * kc.qmin;   // handle_step_rejection! (integrator_utils.jl:135-136)
                else
#endif
                reject_controller(c, kc);
            }
        }
        c.iter++;
        fix_dt_at_bounds(c, P.dtmin);
        modify_dt_for_tstops(c);
        int bad = 0;
        for (int i = tid; i < n; i += NT) bad |= !isfinite(u[i]);
        const bool finite = !__syncthreads_or(bad);
        rc = check_error(c, P, finite);
        if (rc != RC_SUCCESS) break;
        if (c.next_step_tstop && fabs(c.dt) < jl_eps(fabs(c.t))) {
            c.accept_step = 1;
        } else {
            // ---- perform_step! (rosenbrock_perform_step.jl:729-847) ----
            const double t = c.t, dt = c.dt;
            const double idt = 1.0 / dt;
            const double dtgamma = dt * RodasTab::gamma;
            const double sigma = 1.0 / dtgamma;
            odemi_f_wg<double>(fsal, uprev, pp, t, n, tid, NT);
            c.nf += 1;
            const bool new_jac = (c.iter <= 1) ? true : !(c.iter > 1 && c.EEst > 1.0);
            if constexpr (!GJ)
                for (int e = tid; e < n * ld; e += NT) W[e] = (F)0;
            if (sizeof(F) != sizeof(double))
                for (int i = tid; i < n; i += NT) S.fsc[i] = (float)uprev[i];
            __syncthreads();
            if constexpr (GJ) {
#if ODEMI_GJ
                odemi_jac_wg<float>(Jg, ODEMI_GJ_LD, S.fsc, pf, (float)t, n, tid, NT);
#endif
            } else if constexpr (sizeof(F) == sizeof(double)) {
                odemi_jac_wg<double>((double*)W, ld, uprev, pp, t, n, tid, NT);
            } else {
                odemi_jac_wg<float>((float*)W, ld, S.fsc, pf, (float)t, n, tid, NT);
            }
            if (new_jac) {
                c.njacs += 1;
                for (int i = tid; i < n; i += NT) dT[i] = 0.0;
            }
            __syncthreads();
            if constexpr (!GJ)
                for (int i = tid; i < n; i += NT)
                    W[i * ld + i] = (F)fma(-1.0, sigma, (double)W[i * ld + i]);
            if (new_jac) {
                odemi_tgrad_wg<double>(dT, uprev, pp, t, n, tid, NT);
                if (!ODEMI_HAS_TGRAD) c.nf += 1;
            }
            c.nw += 1;
            __syncthreads();
            const double dtd0 = dt * kRd[0];
            for (int i = tid; i < n; i += NT) ks[i] = fsal[i] + dtd0 * dT[i];
            // rb = fsal, r = du, d = u are free during the stage-1 solve
            if constexpr (GJ) {
#if ODEMI_GJ
                gj_load<R>(G, Jg, n, sigma);
                gj_invert<R>(G, n, S.perm, S.gjbuf);
                wg_wsolve_gj<R>(G, S, n, ks, fsal, du, u, uprev, pp, t, sigma, P.refine_sweeps);
#endif
            } else {
                wg_lu<F>(W, ld, n, S.piv, S.perm, red, S.ired);
                wg_wsolve<F>(S, n, ks, fsal, du, u, uprev, pp, t, sigma, P.refine_sweeps);
            }
            for (int i = tid; i < n; i += NT) ks[i] = -ks[i];
            c.nsolve += 1;
            __syncthreads();
            for (int s = 1; s < 8; s++) {
                for (int i = tid; i < n; i += NT) {
                    double acc = uprev[i];
                    for (int q = 0; q < s; q++) acc = fma(kRA[s][q], ks[q * nmax + i], acc);
                    u[i] = acc;
                }
                __syncthreads();
                odemi_f_wg<double>(du, u, pp, fma(kRc[s], dt, t), n, tid, NT);
                c.nf += 1;
                __syncthreads();
                const double dtds = dt * kRd[s];
                double* x = ks + s * nmax;
                for (int i = tid; i < n; i += NT) {
                    double acc = fma(dtds, dT[i], du[i]);
                    for (int q = 0; q < s; q++) acc = fma(kRC[s][q] * idt, ks[q * nmax + i], acc);
                    x[i] = acc;
                }
                __syncthreads();
                if constexpr (GJ) {
#if ODEMI_GJ
                    wg_wsolve_gj<R>(G, S, n, x, fsal, du, u, uprev, pp, t, sigma, P.refine_sweeps);
#endif
                } else {
                    wg_wsolve<F>(S, n, x, fsal, du, u, uprev, pp, t, sigma, P.refine_sweeps);
                }
                for (int i = tid; i < n; i += NT) x[i] = -x[i];
                c.nsolve += 1;
                __syncthreads();
            }
            for (int i = tid; i < n; i += NT) {
                double acc = uprev[i];
                for (int q = 0; q < 7; q++) acc = fma(kRA[7][q], ks[q * nmax + i], acc);
                acc = fma(1.0, ks[7 * nmax + i], acc);
                u[i] = acc;
                double e = 0.0;
                for (int q = 0; q < 7; q++) e = fma(0.0, ks[q * nmax + i], e);
                e = fma(1.0, ks[7 * nmax + i], e);
                du[i] = e;
            }
            if ((SAVES && P.nsaveat > 0) || ODEMI_WG_EVENTS) {   // the event search interpolates every step
                for (int r = 0; r < 3; r++)
                    for (int i = tid; i < n; i += NT) {
                        double acc = 0.0;
                        for (int q = 0; q < 8; q++) acc = fma(kRH[r][q], ks[q * nmax + i], acc);
                        dense[r * nmax + i] = acc;
                    }
            }
            __syncthreads();
            // error norm: the scaled residuals in parallel, then one thread's
            // index-order fma chain (the same value as the sequential norm)
            for (int i = tid; i < n; i += NT)
                fsal[i] = du[i] / fma(max_fast(fabs(uprev[i]), fabs(u[i])), P.reltol, P.abstol);
            __syncthreads();
            if (tid == 0) {
                double sq = 0.0;
                for (int i = 0; i < n; i++) sq = fma(fsal[i], fsal[i], sq);
                red[24] = sqrt(sq / (double)(n > 1 ? n : 1));
            }
            __syncthreads();
            c.EEst = red[24];
        }
        // ---- loopfooter! ----
        const double ttmp = c.t + c.dt;
        const double q = stepsize_controller(c, kc);
#ifdef ODEMI_HAS_ISOUT_WG
        // integrator_utils.jl:612-614, one thread on the new state in LDS
        if (tid == 0) red[25] = odemi_isout_wg<double>(u, pp, ttmp, n) ? 1.0 : 0.0;
        __syncthreads();
        c.isout = (red[25] != 0.0) ? 1 : 0;
#endif
        c.accept_step = !c.isout && (c.EEst <= 1.0);
        if (c.accept_step) {
            c.naccept++;
            c.tprev = c.t;
            if (c.next_step_tstop) { c.dt = c.dtpropose; c.next_step_tstop = 0; c.t = c.tnext; }
            else c.t = ttmp;
            // savevalues!: saveat points up to t, save_everystep; true when it
            // saved at exactly t
            auto savevalues = [&]() -> bool {
                bool exact = false;
                if (!SAVES) return exact;
                while (c.spos != c.send && c.tdir * P.saveat[c.spos] <= c.tdir * c.t) {
                    const double curt = P.saveat[c.spos];
                    c.spos += (c.tdir > 0) ? 1 : -1;
                    if (curt != c.t) {
                        const double th = (curt - c.tprev) / c.dt, th1 = 1.0 - th;
                        const int sidx = c.nsaved;
                        if (sidx < P.max_saves) {
                            if (tid == 0) P.t_out[(i64)sidx * P.N + j] = curt;
                            for (int i = tid; i < n; i += NT) {
                                double in1 = fma(th, dense[2 * nmax + i], dense[nmax + i]);
                                double in2 = fma(th, in1, dense[i]);
                                double in3 = fma(th1, in2, u[i]);
                                P.u_out[((i64)sidx * P.nmax + i) * P.N + j] = fma(th, in3, th1 * uprev[i]);
                            }
                        }
                        c.nsaved = sidx + 1;
                        tlast = curt;
                    } else {
                        if (curt == c.tf && !P.save_end) continue;
                        wg_save_point(c, tlast, P, j, c.t, u, n);
                        exact = true;
                    }
                }
                if (P.save_everystep) {
                    if (c.nsaved == 0 || ((c.t != tlast || c.dt == 0.0) && (P.save_end || c.t != c.tf))) {
                        wg_save_point(c, tlast, P, j, c.t, u, n);
                        exact = true;
                    }
                }
                return exact;
            };
#ifdef ODEMI_WG_CALLBACKS
            // the lane engines' order: controller, callbacks (which run
            // savevalues!), then the tstop pop (an event can move t back)
            accept_controller(c, kc, q, P.dtmin);
#if ODEMI_GJ
            const bool term = wg_handle_callbacks<F, GJ, R>(S, P, j, n, c, tlast, G, Jg, savevalues);
#else
            int G0 = 0;
            const bool term = wg_handle_callbacks<F, GJ, R>(S, P, j, n, c, tlast, G0, Jg, savevalues);
#endif
            handle_tstop(c, P);
            if (term) { rc = RC_TERMINATED; break; }
#else
            handle_tstop(c, P);
            accept_controller(c, kc, q, P.dtmin);
            savevalues();
#endif
        } else {
            c.nreject++;
        }
    }
    return rc;
}

template <bool SAVES>
__device__ __forceinline__ void wg_finish(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double tlast,
                                          int rc) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const int nmax = P.nmax;
    const double* u = S.u;
    // ---- postamble! / result write-out ----
    if ((rc == RC_SUCCESS || rc == RC_TERMINATED) && P.save_end) {   // postamble! runs after terminate! too
        bool doit = (c.nsaved == 0);
        if (!doit)
            doit = (tlast != c.t) && (P.save_end_user == 1 || c.tf_in_saveat || c.t == c.tf || P.nsaveat == 0);
        if (doit) wg_save_point(c, tlast, P, j, c.t, u, n);
    }
    if (P.u_final)
        for (int i = tid; i < nmax; i += NT) P.u_final[(i64)i * P.N + j] = (i < n) ? u[i] : 0.0;
    if (tid == 0) {
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
}

// wg_run with the Gauss-Jordan tile of this trajectory, R = ceil(n / 16), so a
// short chain in a batch padded to nmax does not pay for the largest tile
// (measured: n <= 64 chains 1.55x faster at R = 4 than padded to R = 8).
// Only the tile sizes in ODEMI_GJ_RMASK are instantiated; the largest always.
template <typename F, bool SAVES, bool GJ, int RR = 1>
__device__ __forceinline__ int wg_run_tiled(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c, double& tlast,
                                            float* Jg, int limit) {
#if ODEMI_GJ
    if constexpr (GJ) {
        if constexpr (RR < ODEMI_GJ_R) {
            if constexpr (((ODEMI_GJ_RMASK >> (RR - 1)) & 1) != 0)
                if (((n + 15) >> 4) == RR) return wg_run<F, SAVES, GJ, RR>(S, P, j, n, c, tlast, Jg, limit);
            return wg_run_tiled<F, SAVES, GJ, RR + 1>(S, P, j, n, c, tlast, Jg, limit);
        } else {
            return wg_run<F, SAVES, GJ, ODEMI_GJ_R>(S, P, j, n, c, tlast, Jg, limit);
        }
    }
#endif
    return wg_run<F, SAVES, GJ>(S, P, j, n, c, tlast, Jg, limit);
}

template <typename F, bool SAVES, bool GJ>
__device__ __forceinline__ void rodas_wg_impl(const Params& P) {
    static_assert(!GJ || (sizeof(F) == 4 && ODEMI_GJ), "Gauss-Jordan variant: FP32, n <= 128");
    extern __shared__ __attribute__((aligned(16))) unsigned char odemi_lds[];
    const int tid = threadIdx.x;
    const WGPtrs S = wg_layout<F>(odemi_lds, P.nmax, SAVES || ODEMI_WG_EVENTS, !GJ);
    float* Jg = nullptr;
#if ODEMI_GJ
    if constexpr (GJ) Jg = P.wg_jbuf + (i64)blockIdx.x * ODEMI_GJ_LD * ODEMI_GJ_LD;
#endif
    for (;;) {
        if (tid == 0) ((i64*)S.red)[30] = (i64)atomicAdd(&P.counters[0], 1ULL);
        __syncthreads();
        const i64 slot = ((i64*)S.red)[30];
        __syncthreads();
        if (slot >= P.N) break;
        const i64 j = P.order ? (i64)P.order[slot] : slot;
        const int n = P.nvec ? P.nvec[j] : P.nmax;
        Ctl c;
        double tlast;
        wg_begin<GJ>(S, P, j, n, c, tlast, Jg);
        const int rc = wg_run_tiled<F, SAVES, GJ>(S, P, j, n, c, tlast, Jg, 0x7fffffff);
        wg_finish<SAVES>(S, P, j, n, c, tlast, rc);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Serving (continuous batching) with the workgroup kernel: the slot pool keeps
// the controller in the plain-engine layout and the rows u, uprev, dT in
// st_vec; W, the stages and the Jacobian are rebuilt by every attempt, so
// that is the whole state.  Slot state st_int row 7: 0 free, 1 active.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wg_store_state(const WGPtrs& S, const Params& P, i64 j, int n, const Ctl& c,
                                               int active) {
    const i64 M = P.N;
    if (threadIdx.x == 0) {
        plain_store_ctl<1>(c, P, j, active);
        P.st_int[9 * M + j] = c.njacs; P.st_int[10 * M + j] = c.nw; P.st_int[11 * M + j] = c.nsolve;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        P.st_vec[((i64)0 * P.nmax + i) * M + j] = S.u[i];
        P.st_vec[((i64)1 * P.nmax + i) * M + j] = S.uprev[i];
        P.st_vec[((i64)2 * P.nmax + i) * M + j] = S.dT[i];
    }
}
__device__ __forceinline__ void wg_load_state(const WGPtrs& S, const Params& P, i64 j, int n, Ctl& c) {
    const i64 M = P.N;
    plain_load_ctl<1>(c, P, j);
    c.njacs = P.st_int[9 * M + j]; c.nw = P.st_int[10 * M + j]; c.nsolve = P.st_int[11 * M + j];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        S.u[i] = P.st_vec[((i64)0 * P.nmax + i) * M + j];
        S.uprev[i] = P.st_vec[((i64)1 * P.nmax + i) * M + j];
        S.dT[i] = P.st_vec[((i64)2 * P.nmax + i) * M + j];
    }
}

// Admit slots [slot0, slot0 + nslots): one workgroup per slot.
template <typename F, bool GJ>
__device__ __forceinline__ void rodas_wg_serve_init(const Params& P) {
    extern __shared__ __attribute__((aligned(16))) unsigned char odemi_lds[];
    const WGPtrs S = wg_layout<F>(odemi_lds, P.nmax, ODEMI_WG_EVENTS != 0, !GJ);
    float* Jg = nullptr;
#if ODEMI_GJ
    if constexpr (GJ) Jg = P.wg_jbuf + (i64)blockIdx.x * ODEMI_GJ_LD * ODEMI_GJ_LD;
#endif
    for (i64 q = blockIdx.x; q < P.nslots; q += gridDim.x) {
        const i64 j = P.slot0 + q;
        const int n = P.nvec[j];
        const Params Q = slot_params(P, j);
        Ctl c;
        double tlast;
        wg_begin<GJ>(S, Q, j, n, c, tlast, Jg);
        wg_store_state(S, Q, j, n, c, 1);
        __syncthreads();
    }
}

// Advance every active slot in [0, nslots) by up to tick_attempts attempts;
// slots are handed out through counters[0] (cleared by the host per tick).
template <typename F, bool GJ>
__device__ __forceinline__ void rodas_wg_serve_tick(const Params& P) {
    extern __shared__ __attribute__((aligned(16))) unsigned char odemi_lds[];
    const int tid = threadIdx.x;
    const WGPtrs S = wg_layout<F>(odemi_lds, P.nmax, ODEMI_WG_EVENTS != 0, !GJ);
    float* Jg = nullptr;
#if ODEMI_GJ
    if constexpr (GJ) Jg = P.wg_jbuf + (i64)blockIdx.x * ODEMI_GJ_LD * ODEMI_GJ_LD;
#endif
    for (;;) {
        if (tid == 0) ((i64*)S.red)[30] = (i64)atomicAdd(&P.counters[0], 1ULL);
        __syncthreads();
        const i64 j = ((i64*)S.red)[30];
        __syncthreads();
        if (j >= P.nslots) break;
        if (P.st_int[7 * P.N + j] != 1) continue;
        const int n = P.nvec[j];
        const Params Q = slot_params(P, j);
        wg_load_params(S, Q, j);
        Ctl c;
        wg_load_state(S, Q, j, n, c);
#if ODEMI_GJ
        if constexpr (GJ)
            for (int e = tid; e < n * ODEMI_GJ_LD; e += blockDim.x) Jg[e] = 0.0f;
#endif
        __syncthreads();
        double tlast = c.t;
        const int rc = wg_run_tiled<F, false, GJ>(S, Q, j, n, c, tlast, Jg, P.tick_attempts);
        if (rc) {
            wg_finish<false>(S, Q, j, n, c, tlast, rc);
            if (tid == 0) {
                P.st_int[7 * P.N + j] = 0;
                atomicAdd(&P.req_done[P.slot_req[j]], 1ULL);
            }
        } else {
            wg_store_state(S, Q, j, n, c, 1);
        }
        __syncthreads();
    }
}

}  // namespace odemi

#endif  // ODEMI_WITH_WG
