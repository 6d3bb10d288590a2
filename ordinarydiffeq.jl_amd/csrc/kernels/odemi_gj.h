// odemi_gj.h -- FP32 Gauss-Jordan inverse of W = J - sigma I, held in the
// registers of the 256-thread workgroup of the mixed workgroup Rodas5P kernels
// (n <= 128, one trajectory per workgroup), and the stage solve with it: each
// stage solve is one parallel matrix-vector product instead of two serial
// substitutions.  A trajectory of n states uses the tile R = ceil(n / 16)
// (NP = 16 R >= n; rows and columns in [n, NP) are identity padding).
//
// Two forms with the same result semantics (pivot order, W^{-1} b = z[pk],
// z = T b[pk]):
//   ODEMI_GJ_MFMA 1 (default): blocked by 16 columns, each panel factored by
//     one wave and the rest of the matrix updated by rank-16 products on the
//     matrix cores (v_mfma_f32_16x16x4f32);
//   ODEMI_GJ_MFMA 0: one rank-1 update per step on the vector ALUs, each
//     thread holding an R x R tile.
#pragma once
#if ODEMI_N <= 128
#define ODEMI_GJ 1
#define ODEMI_GJ_R ((ODEMI_N + 15) / 16)   // tile of the largest state (nmax)
// Tile sizes R = ceil(n/16) present in the batch (bit R-1); the problem source
// may define it from its state sizes, else every R up to ODEMI_GJ_R.
#ifndef ODEMI_GJ_RMASK
#define ODEMI_GJ_RMASK 0xff
#endif
#ifndef ODEMI_GJ_MFMA
#define ODEMI_GJ_MFMA 1
#endif
// Row stride of each workgroup's FP32 Jacobian block, which holds LD x LD
// floats (a compile-time stride: immediate load offsets).
#define ODEMI_GJ_LD (16 * ODEMI_GJ_R)

namespace odemi {


// Sum over the 16 lanes of each DPP row; the total lands in the row's lane 15.
// Fixed addition order (inclusive Hillis-Steele scan), so results are
// deterministic.  All 64 lanes must be active.
__device__ __forceinline__ float dpp_row_sum16(float x) {
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, false));  // row_shr:1
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xf, 0xf, false));  // row_shr:2
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xf, 0xf, false));  // row_shr:4
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xf, 0xf, false));  // row_shr:8
    return x;
}
// Maximum over the wave, returned to every lane.  All 64 lanes must be active.
// Lanes without a source read 0, the identity of the unsigned maximum.
__device__ __forceinline__ unsigned dpp_wave_max_u32(unsigned x) {
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, true));  // row_bcast:15
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, true));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

__device__ __forceinline__ unsigned gj_key(float v, int i) {
    // larger |v| wins, ties to the smaller row; a NaN entry competes as zero so
    // the pivot is always an unused row < n (n <= 128)
    const float a = fabsf(v);
    return ((a != a) ? 0u : (__float_as_uint(a) & ~127u)) | (unsigned)(127 - i);
}

#if ODEMI_GJ_MFMA
// Blocked form.  The matrix (NP = 16 R rows and columns, identity padding
// outside the n x n block) is a grid of R x R blocks of 16 x 16, each in the
// accumulator layout of v_mfma_f32_16x16x4f32: lane l of a wave holds rows
// 4 (l / 16) + r (r < 4) of column l % 16 of the block.  Wave w holds block
// rows w, w + 4 (D[m][J] is block (w + 4 m, J)).
typedef float gjf4 __attribute__((ext_vector_type(4)));

template <int R>
struct GJTile {
    static constexpr int MB = (R + 3) / 4;   // block rows per wave
    gjf4 D[MB][R];
};

// Offset (floats) of the matrix-vector result in the Gauss-Jordan LDS buffer
// of tile R; the buffer holds 51 NP + 16 floats:
//   pan [2][NP][16]  the panel, then its multipliers (alternating by panel parity)
//   wp  [16][NP]     the panel's pivot rows before the panel's update
//   scl [NP]         row scales
//   pst int[NP]      step at which each row became a pivot row (-1: not yet)
//   zbuf [NP]        matrix-vector result
//   prow [16]        the current step's pivot row (panel wave)
template <int R>
__device__ __forceinline__ constexpr int gj_zbuf_off() { return 50 * 16 * R; }

// T <- W = J - sigma I for the n x n leading block of J (FP32, row stride
// ODEMI_GJ_LD), identity outside it.  The diagonal shift is formed in FP64 and
// rounded once.  (Loading the whole tile unconditionally and selecting the
// padding afterwards measured 4 % slower on uneven chains: short chains then
// load their full padded tile.)
template <int R>
__device__ __forceinline__ void gj_load(GJTile<R>& G, const float* J, int n, double sigma) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
    for (int m = 0; m < GJTile<R>::MB; m++) {
        const int I = w + 4 * m;
#pragma unroll
        for (int b = 0; b < R; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * I + 4 * g + r, j = 16 * b + c;
                float v = (i == j) ? 1.0f : 0.0f;
                if (i < n && j < n) {
                    v = J[i * ODEMI_GJ_LD + j];
                    if (i == j) v = (float)fma(-1.0, sigma, (double)v);
                }
                G.D[m][b][r] = v;
            }
    }
}

// In-place Gauss-Jordan with partial pivoting and no row interchanges: step k
// pivots on the largest |W[i][k]| among rows not yet used (ties to the smaller
// row) and exchanges the roles of row p_k and column k.  On return
// W^{-1} b = z[pk], z = T * b[pk] (pk[k] = p_k).
//
// Rows are held raw with W = diag(sc) Raw: a pivot row is not rescaled, its
// scale becomes d = 1/W[p][k]; the update of any other row is proportional to
// that row's own column-k entry, so its scale factors out and every step is
// Raw[i][j] -= g_i Raw[p][j] (g_i = Raw[i][k] d, g_p = 0), column k becoming
// -g_i.  The 1 that belongs at (p_k, k) is never read afterwards (row p_k is
// never a pivot row again, column k never a pivot column again), so the ones
// and the scales are applied at the end.
//
// Blocked by 16 columns: the 16 steps of a panel, restricted to the panel's
// columns, are done by one wave (wave-level pivot search, no workgroup
// barrier); their product is I + C S, S selecting the 16 pivot rows and C the
// panel's final values less the implied ones, so the other columns are brought
// up to date at once, Raw[:, rest] += C Raw_before[P, rest], as 4 MFMAs per
// 16 x 16 block (the panel's own block is updated too and then replaced, so
// the products run without branches).  Three workgroup barriers per panel.
template <int R>
__device__ __forceinline__ void gj_invert(GJTile<R>& G, int n, int* pk, float* gjbuf) {
    constexpr int NP = 16 * R, NQ = (NP + 63) / 64, MB = GJTile<R>::MB;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    float* pan0 = gjbuf;
    float* wp = gjbuf + 32 * NP;
    float* scl = gjbuf + 48 * NP;
    int* pst = (int*)(gjbuf + 49 * NP);
    float* prowb = gjbuf + 51 * NP;
    for (int i = tid; i < NP; i += blockDim.x) pst[i] = -1;
    // panel 0
#pragma unroll
    for (int m = 0; m < MB; m++)
        if (w + 4 * m < R) {
#pragma unroll
            for (int r = 0; r < 4; r++) pan0[(16 * (w + 4 * m) + 4 * g + r) * 16 + c] = G.D[m][0][r];
        }
    __syncthreads();
    for (int b = 0; b < R; b++) {
        const int k0 = 16 * b;
        if (k0 >= n) break;
        const int mc = min(16, n - k0);
        float* pan = pan0 + (b & 1) * 16 * NP;
        // ---- the panel's steps, by one wave ----
        if (w == (b & 3)) {
            float pv[NQ][16];
            unsigned used = 0;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int i = l + 64 * q;
                if (i < NP) {
#pragma unroll
                    for (int v4 = 0; v4 < 4; v4++) {
                        const float4 t4 = *(const float4*)(pan + i * 16 + 4 * v4);
                        pv[q][4 * v4] = t4.x; pv[q][4 * v4 + 1] = t4.y;
                        pv[q][4 * v4 + 2] = t4.z; pv[q][4 * v4 + 3] = t4.w;
                    }
                    if (pst[i] >= 0) used |= 1u << q;
                } else {
#pragma unroll
                    for (int cc = 0; cc < 16; cc++) pv[q][cc] = 0.0f;
                }
            }
#pragma unroll
            for (int kt = 0; kt < 16; kt++) {
                if (kt >= mc) continue;   // uniform
                unsigned key = 0;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int i = l + 64 * q;
                    if (i < n && !((used >> q) & 1u)) key = max(key, gj_key(pv[q][kt], i));
                }
                key = dpp_wave_max_u32(key);
                const int p = 127 - (int)(key & 127u);
                // the pivot row through LDS: its lane writes it and every lane
                // reads it back after a wave barrier (without the barrier the
                // compiler may order a lane's read before another lane's write)
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    if (l + 64 * q == p) {
#pragma unroll
                        for (int v4 = 0; v4 < 4; v4++)
                            *(float4*)(prowb + 4 * v4) =
                                make_float4(pv[q][4 * v4], pv[q][4 * v4 + 1], pv[q][4 * v4 + 2], pv[q][4 * v4 + 3]);
                    }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                float prow[16];
#pragma unroll
                for (int v4 = 0; v4 < 4; v4++) {
                    const float4 t4 = *(const float4*)(prowb + 4 * v4);
                    prow[4 * v4] = t4.x; prow[4 * v4 + 1] = t4.y; prow[4 * v4 + 2] = t4.z; prow[4 * v4 + 3] = t4.w;
                }
                const float d = __builtin_amdgcn_rcpf(prow[kt]);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int i = l + 64 * q;
                    const float gq = (i == p) ? 0.0f : pv[q][kt] * d;
#pragma unroll
                    for (int cc = 0; cc < 16; cc++)
                        if (cc != kt) pv[q][cc] = fmaf(-gq, prow[cc], pv[q][cc]);
                    pv[q][kt] = -gq;
                    if (i == p) used |= 1u << q;
                }
                if (l == 0) {
                    pk[k0 + kt] = p;
                    scl[p] = d;
                    pst[p] = k0 + kt;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int i = l + 64 * q;
                if (i < NP) {
#pragma unroll
                    for (int v4 = 0; v4 < 4; v4++)
                        *(float4*)(pan + i * 16 + 4 * v4) =
                            make_float4(pv[q][4 * v4], pv[q][4 * v4 + 1], pv[q][4 * v4 + 2], pv[q][4 * v4 + 3]);
                }
            }
        }
        __syncthreads();
        // ---- the panel's pivot rows as they were before its update, by
        // their owners; rows of wp past the panel's last step are zero ----
#pragma unroll
        for (int m = 0; m < MB; m++)
            if (w + 4 * m < R) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int mm = pst[16 * (w + 4 * m) + 4 * g + r] - k0;
                    if (mm >= 0 && mm < 16) {
#pragma unroll
                        for (int J = 0; J < R; J++) wp[mm * NP + 16 * J + c] = G.D[m][J][r];
                    }
                }
            }
        for (int e = mc * NP + tid; e < 16 * NP; e += blockDim.x) wp[e] = 0.0f;
        __syncthreads();
        // ---- rank-16 update of every column block (4 MFMAs per block, the
        // blocks interleaved); then the panel's block takes the multipliers
        // and the next panel is published ----
#pragma unroll
        for (int m = 0; m < MB; m++) {
            const int I = w + 4 * m;
            if (I < R) {
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    const float a = pan[(16 * I + c) * 16 + 4 * kc + g];
                    float bv[R];
#pragma unroll
                    for (int J = 0; J < R; J++) bv[J] = wp[(4 * kc + g) * NP + 16 * J + c];
#pragma unroll
                    for (int J = 0; J < R; J++)
                        G.D[m][J] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv[J], G.D[m][J], 0, 0, 0);
                }
#pragma unroll
                for (int J = 0; J < R; J++)
                    if (J == b) {
#pragma unroll
                        for (int r = 0; r < 4; r++) G.D[m][J][r] = pan[(16 * I + 4 * g + r) * 16 + c];
                    }
            }
        }
        if (k0 + 16 < n) {
            float* pn = pan0 + ((b + 1) & 1) * 16 * NP;
#pragma unroll
            for (int m = 0; m < MB; m++)
                if (w + 4 * m < R) {
#pragma unroll
                    for (int J = 0; J < R; J++)
                        if (J == b + 1) {
#pragma unroll
                            for (int r = 0; r < 4; r++) pn[(16 * (w + 4 * m) + 4 * g + r) * 16 + c] = G.D[m][J][r];
                        }
                }
        }
        __syncthreads();
    }
    // the implied ones, then the row scales
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int I = w + 4 * m;
        if (I < R) {
#pragma unroll
            for (int J = 0; J < R; J++) {
                const int j = 16 * J + c;
                const int pj = (j < n) ? pk[j] : -1;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 16 * I + 4 * g + r;
                    float v = G.D[m][J][r];
                    if (pj == i) v += 1.0f;
                    G.D[m][J][r] = v * ((i < n) ? scl[i] : 1.0f);
                }
            }
        }
    }
}

// x <- W^{-1} x (x FP64 in LDS, accumulated in FP32).  All threads must call.
template <int R>
__device__ __forceinline__ void gj_solve(const GJTile<R>& G, int n, const int* pk, double* x, float* zbuf) {
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    float bp[R];
#pragma unroll
    for (int J = 0; J < R; J++) {
        const int j = 16 * J + c;
        bp[J] = (j < n) ? (float)x[pk[j]] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < GJTile<R>::MB; m++) {
        const int I = w + 4 * m;
        if (I < R) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float z = 0.0f;
#pragma unroll
                for (int J = 0; J < R; J++) z = fmaf(G.D[m][J][r], bp[J], z);
                z = dpp_row_sum16(z);
                if (c == 15) zbuf[16 * I + 4 * g + r] = z;
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < n; k += blockDim.x) x[k] = (double)zbuf[pk[k]];
    __syncthreads();
}
#else
// R x R tile per thread; a trajectory of n states uses R = ceil(n / 16)
// (16 R >= n; rows and columns in [n, 16 R) are identity padding).
template <int R>
struct GJTile {
    float T[R][R];
};

// T <- W = J - sigma I for the n x n leading block of J (FP32, row stride
// ODEMI_GJ_LD), identity outside it.
template <int R>
__device__ __forceinline__ void gj_load(GJTile<R>& G, const float* J, int n, double sigma) {
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
    for (int a = 0; a < R; a++)
#pragma unroll
        for (int b = 0; b < R; b++) {
            const int i = ti + 16 * a, j = tj + 16 * b;
            float v = (i == j) ? 1.0f : 0.0f;
            if (i < n && j < n) {
                v = J[i * ODEMI_GJ_LD + j];
                if (i == j) v = (float)fma(-1.0, sigma, (double)v);
            }
            G.T[a][b] = v;
        }
}

// In-place Gauss-Jordan with partial pivoting and no row interchanges: step k
// pivots on the largest |W[i][k]| among rows not yet used (ties to the smaller
// row) and exchanges the roles of row p_k and column k.  On return
// W^{-1} b = z[pk], z = T * b[pk] (pk[k] = p_k).
//
// The tile holds raw rows R with W = diag(sc) R.  A pivot row is not rescaled
// in the tile: its scale sc_p becomes d = 1/W[p][k] and its raw entries stay,
// so the only per-element work of a step is the rank-1 update of the other
// rows, and no thread has to write a register chosen by a runtime index.
// Rows not yet used keep sc = 1, so the pivot search and the pivot row read
// raw values.  The scales (one LDS word per row) are folded into the tile at
// the end.
//
// The step is one fma per element with no selects: column k is published with
// 0 in the pivot row's slot (so the pivot row's update is a no-op), the
// threads holding column k update it from zero with d in place of r[k] d (so
// it becomes -c d), and the 1 that belongs at (p_k, k) is added at the end,
// before the scales.  That entry is only ever updated afterwards, never read
// as a pivot column or pivot row entry, so adding it last is exact algebra.
// The pivot itself is read from the staged pivot row.
//
// One barrier per step.  Before the barrier that opens step k the owners
// publish column k, column k+1 and (all R row slots of the thread row holding)
// row p_k, as step k-1 left them.  After its update every wave forms column
// k+1 as step k leaves it (the same fma as the owners' update, so
// bit-identical) and searches it for p_{k+1}, so the owners of row p_{k+1}
// can stage it before the next barrier.  Buffers alternate by step parity.
//   cbuf, nbuf [2][16][R]: entry (ti, a) is row ti + 16 a
//   stg [2][R][16][R]:     entry (a, tj, b) is T[a][b] of thread (p_k % 16, tj)
// Offset (floats) of the matrix-vector result in the Gauss-Jordan LDS buffer
// of tile R: after the column, next-column and row-staging buffers.
template <int R>
__device__ __forceinline__ constexpr int gj_zbuf_off() { return 5 * 16 * R + 2 * R * 16 * R; }

template <int R>
__device__ __forceinline__ int gj_cidx(int i) { return (i & 15) * R + (i >> 4); }

template <int R>
__device__ __forceinline__ void gj_invert(GJTile<R>& G, int n, int* pk, float* gjbuf) {
    constexpr int NP = 16 * R, NQ = (NP + 63) / 64;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15, lane = tid & 63;
    float* cbuf = gjbuf;
    float* nbuf = gjbuf + 2 * NP;
    float* stg = gjbuf + 4 * NP;
    float* scl = gjbuf + 4 * NP + 2 * R * NP;   // [NP] scale of each pivot row
    unsigned used = 0;   // bit q: row lane + 64 q has been a pivot row
    // ---- prologue: columns 0 and 1, pivot p_0, row p_0 ----
    if (tj == 0) {
#pragma unroll
        for (int a = 0; a < R; a++) cbuf[ti * R + a] = G.T[a][0];
    }
    if (tj == 1 && n > 1) {
#pragma unroll
        for (int a = 0; a < R; a++) nbuf[ti * R + a] = G.T[a][0];
    }
    __syncthreads();
    int p;
    {
        unsigned key = 0;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int i = lane + 64 * q;
            if (i < n) key = max(key, gj_key(cbuf[gj_cidx<R>(i)], i));
        }
        key = dpp_wave_max_u32(key);
        p = 127 - (int)(key & 127u);
    }
    if (tid == 0) pk[0] = p;
    if (ti == (p & 15)) {
#pragma unroll
        for (int a = 0; a < R; a++)
#pragma unroll
            for (int b = 0; b < R; b++) stg[(a * 16 + tj) * R + b] = G.T[a][b];
    }
    __syncthreads();
    // every wave has searched column 0: its pivot slot can go to zero
    if (tj == 0 && ti == (p & 15)) cbuf[ti * R + (p >> 4)] = 0.0f;
    __syncthreads();
    // ---- steps ----
#pragma unroll
    for (int kb = 0; kb < R; kb++) {
        const int kend = min(16, n - 16 * kb);
        for (int kt = 0; kt < kend; kt++) {
            const int k = 16 * kb + kt, k1 = k + 1, k2 = k + 2;
            const int par = k & 1;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (p == lane + 64 * q) used |= 1u << q;
            const float* cb = cbuf + par * NP;
            const float* nb = nbuf + par * NP;
            const float* st = stg + par * R * NP;
            const int pa = p >> 4;
            const bool colk = (tj == kt);
            // all reads of the step first (the search for p_{k+1} only needs
            // these, so its latency overlaps the tile update).  At the last step
            // k1 = n <= NP: the reads stay inside the buffers and are unused.
            const float piv = st[(pa * 16 + kt) * R + kb];
            float c[R], r[R], cq[NQ], nq[NQ];
#pragma unroll
            for (int a = 0; a < R; a++) c[a] = cb[ti * R + a];
#pragma unroll
            for (int b = 0; b < R; b++) r[b] = st[(pa * 16 + tj) * R + b];
            const float r1 = st[(pa * 16 + (k1 & 15)) * R + (k1 >> 4)];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int ci = gj_cidx<R>(min(lane + 64 * q, NP - 1));
                cq[q] = cb[ci];
                nq[q] = nb[ci];
            }
            const float d = __builtin_amdgcn_rcpf(piv);
            // pivot of step k+1 from column k+1 as step k leaves it (same fma as
            // the owners' update)
            const float rs1 = r1 * d;
            unsigned key = 0;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int i = lane + 64 * q;
                if (i < n && !((used >> q) & 1u)) key = max(key, gj_key(fmaf(-cq[q], rs1, nq[q]), i));
            }
            // rank-1 update (c is 0 on row p_k); column k becomes 0 - c d
            float rd[R];
#pragma unroll
            for (int b = 0; b < R; b++) rd[b] = r[b] * d;
            rd[kb] = colk ? d : rd[kb];
#pragma unroll
            for (int a = 0; a < R; a++)
#pragma unroll
                for (int b = 0; b < R; b++) {
                    const float t = (b == kb && colk) ? 0.0f : G.T[a][b];
                    G.T[a][b] = fmaf(-c[a], rd[b], t);
                }
            if (tid == 0) scl[p] = d;
            key = dpp_wave_max_u32(key);
            if (k1 >= n) break;
            p = 127 - (int)(key & 127u);
            if (tid == 0) pk[k1] = p;
            // publish column k+1, column k+2 and row p_{k+1} for the next step
            float* cbn = cbuf + (NP - par * NP);
            float* nbn = nbuf + (NP - par * NP);
            float* stn = stg + (R * NP - par * R * NP);
            if (tj == (k1 & 15)) {
                if (kt < 15) {
#pragma unroll
                    for (int a = 0; a < R; a++) cbn[ti * R + a] = G.T[a][kb];
                } else {
#pragma unroll
                    for (int a = 0; a < R; a++) cbn[ti * R + a] = G.T[a][kb + 1 < R ? kb + 1 : kb];
                }
            }
            if (k2 < n && tj == (k2 & 15)) {
                if (kt < 14) {
#pragma unroll
                    for (int a = 0; a < R; a++) nbn[ti * R + a] = G.T[a][kb];
                } else {
#pragma unroll
                    for (int a = 0; a < R; a++) nbn[ti * R + a] = G.T[a][kb + 1 < R ? kb + 1 : kb];
                }
            }
            if (ti == (p & 15)) {
#pragma unroll
                for (int a = 0; a < R; a++)
#pragma unroll
                    for (int b = 0; b < R; b++) stn[(a * 16 + tj) * R + b] = G.T[a][b];
            }
            // the pivot slot of column k+1 goes to zero (its publisher wrote it
            // above; nobody reads this parity's buffers before the barrier)
            if (tj == (k1 & 15) && ti == (p & 15)) cbn[ti * R + (p >> 4)] = 0.0f;
            __syncthreads();
        }
    }
    __syncthreads();
    float pjv[R];
#pragma unroll
    for (int b = 0; b < R; b++) {
        const int j = tj + 16 * b;
        pjv[b] = (j < n) ? (float)pk[j] : -1.0f;
    }
#pragma unroll
    for (int a = 0; a < R; a++) {
        const int i = ti + 16 * a;
        const float sa = (i < n) ? scl[i] : 1.0f;
#pragma unroll
        for (int b = 0; b < R; b++) {
            const float v = (pjv[b] == (float)i) ? G.T[a][b] + 1.0f : G.T[a][b];
            G.T[a][b] = v * sa;
        }
    }
}

// x <- W^{-1} x (x FP64 in LDS, accumulated in FP32).  All threads must call.
template <int R>
__device__ __forceinline__ void gj_solve(const GJTile<R>& G, int n, const int* pk, double* x, float* zbuf) {
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    float bp[R];
#pragma unroll
    for (int b = 0; b < R; b++) {
        const int j = tj + 16 * b;
        bp[b] = (j < n) ? (float)x[pk[j]] : 0.0f;
    }
#pragma unroll
    for (int a = 0; a < R; a++) {
        float z = 0.0f;
#pragma unroll
        for (int b = 0; b < R; b++) z = fmaf(G.T[a][b], bp[b], z);
        z = dpp_row_sum16(z);
        if (tj == 15) zbuf[ti + 16 * a] = z;
    }
    __syncthreads();
    for (int k = tid; k < n; k += blockDim.x) x[k] = (double)zbuf[pk[k]];
    __syncthreads();
}
#endif  // ODEMI_GJ_MFMA

}  // namespace odemi
#else
#define ODEMI_GJ 0
#endif  // ODEMI_N <= 128
