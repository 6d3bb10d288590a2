// odemi_runtime.cpp -- host runtime of the MI355X ensemble ODE engine.
//
// Implements the C ABI in include/odemi.h: device context, hipRTC compilation
// of traced problems (with an on-disk code-object cache), device workspace
// management and the launch logic of the two engines:
//   plain       one kernel launch per step attempt over all live trajectories,
//               integrator state round-trips through HBM (FP64 baseline);
//   persistent  one launch; each lane (or workgroup) owns a trajectory in
//               registers/LDS, retires it and refills from a work queue.
#include "odemi.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <numeric>
#include <sstream>
#include <string>
#include <climits>
#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace {

const char* kDeviceSource =
#include "odemi_device_src.inc"
    ;

// rows of the int32 integrator state of the plain engine and the serving pool
// (iter, success_iter, accept, next_step_tstop, nf, naccept, nreject, slot
// state, saveat cursor, njacs, nw, nsolve, tstops cursor, isout | reeval_fsal
// << 1, ncondition, last continuous event | its vector components << 8)
constexpr size_t kIntRows = 16;
// DEStats rows written per trajectory: nf naccept nreject njacs nw nsolve ncondition
constexpr size_t kStatRows = 7;
// odemi_features bits: the lane kernels' per-step hooks (isoutofdomain,
// discrete, continuous callbacks); the workgroup kernels' callbacks are bits
// 8 (discrete) and 16 (continuous, whose event search keeps dense rows in LDS)
constexpr int kLaneHooks = 1 | 2 | 4;

struct Params {  // must match odemi::Params in kernels/odemi_device.h
    int64_t N;
    int nmax, np;
    const int* nvec;
    const double *u0, *p, *t0, *tf;
    const int* order;
    double reltol, abstol, dtmax, dtmin, dt_user;
    int64_t maxiters;
    const double* saveat;
    int nsaveat;
    const double* tstops;
    int ntstops;
    int save_start, save_end, save_everystep, save_end_user;
    int max_saves;
    double* t_out;
    double* u_out;
    int* nsaved;
    double* u_final;
    double* t_final;
    int* retcode;
    int* stats;
    unsigned long long* counters;
    double* st_scalar;
    double* st_vec;
    int* st_int;
    int refine_sweeps;
    int tick_attempts;
    int64_t slot0, nslots;
    const int* slot_req;
    unsigned long long* req_done;
    const double* req_saveat;
    const double* req_tstops;
    const int* req_opts;
    int sa_cap, ts_cap;
    float* wg_jbuf;
    double* pre_dt;
    double* pre_k1;
    double* rec_in;
    double* rec_out;
    const int* inv;
};
// Doubles per queue-order record (the device code's rec_in_words /
// rec_out_words for n states and np parameters, np >= 1).
constexpr int rec_in_words(int n, int np) { return (2 * n + np + 3 + 1) & ~1; }
constexpr int rec_out_words(int n) { return (n + 1 + 4 + 1) & ~1; }
// ints per serving request-table entry (kReqOpts): nsaveat, ntstops,
// save_start, save_end, save_end_user
constexpr int kReqOpts = 8;

// Longest-predicted-cost-first work order, key_j = (tf_j - t0_j) * n_j^3, by a
// counting sort over the top 11 bits of the FP32 key (exponent and three
// mantissa bits: monotone for keys >= 0, buckets 1/8 wide in relative cost).
// Inside a bucket the order is arbitrary; a trajectory's result does not depend
// on the lane that runs it.  Three passes: bucket counts (LDS histogram per
// block, one global atomic per non-empty bin), a descending scan of the bins,
// and the scatter (per block one global atomic per non-empty bin reserves a
// range, LDS cursors place the items).
constexpr int kCostBins = 2048;
constexpr int kCostItems = 16;                       // items per thread
constexpr int kCostTile = 256 * kCostItems;          // items per block
__device__ __forceinline__ int cost_bin(const double* t0, const double* tf, const int* nvec, int64_t j) {
    double c = tf[j] - t0[j];
    if (nvec) { double n = (double)nvec[j]; c *= n * n * n; }
    const float k = (c > 0.0) ? (float)c : 0.0f;     // NaN -> 0
    return (int)(__float_as_uint(k) >> 20);          // < 2048 for k in [0, inf]
}
__global__ void __launch_bounds__(256) cost_hist_kernel(const double* t0, const double* tf, const int* nvec,
                                                       int64_t N, unsigned* counts) {
    __shared__ unsigned h[kCostBins];
    for (int b = threadIdx.x; b < kCostBins; b += 256) h[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kCostTile;
    for (int q = 0; q < kCostItems; q++) {
        const int64_t j = base + (int64_t)q * 256 + threadIdx.x;
        if (j < N) atomicAdd(&h[cost_bin(t0, tf, nvec, j)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kCostBins; b += 256)
        if (h[b]) atomicAdd(&counts[b], h[b]);
}
// counts -> start offsets, highest bin first (one block of 256 threads)
__global__ void __launch_bounds__(256) cost_scan_kernel(unsigned* counts) {
    constexpr int per = kCostBins / 256;
    __shared__ unsigned part[256];
    const int t = threadIdx.x;
    unsigned sum = 0;
    for (int q = 0; q < per; q++) sum += counts[kCostBins - 1 - (t * per + q)];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = part[t] - sum;
    for (int q = 0; q < per; q++) {
        const int b = kCostBins - 1 - (t * per + q);
        const unsigned c = counts[b];
        counts[b] = run;
        run += c;
    }
}
__global__ void __launch_bounds__(256) cost_scatter_kernel(const double* t0, const double* tf, const int* nvec,
                                                          int64_t N, unsigned* cursor, int* order, int* inv) {
    __shared__ unsigned h[kCostBins];
    for (int b = threadIdx.x; b < kCostBins; b += 256) h[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kCostTile;
    int bin[kCostItems];
    unsigned rank[kCostItems];
    for (int q = 0; q < kCostItems; q++) {
        const int64_t j = base + (int64_t)q * 256 + threadIdx.x;
        bin[q] = (j < N) ? cost_bin(t0, tf, nvec, j) : -1;
        rank[q] = (bin[q] >= 0) ? atomicAdd(&h[bin[q]], 1u) : 0u;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kCostBins; b += 256)
        if (h[b]) h[b] = atomicAdd(&cursor[b], h[b]);   // this block's range in bin b
    __syncthreads();
    for (int q = 0; q < kCostItems; q++) {
        const int64_t j = base + (int64_t)q * 256 + threadIdx.x;
        if (bin[q] >= 0) {
            const unsigned pos = h[bin[q]] + rank[q];
            order[pos] = (int)j;
            if (inv) inv[j] = (int)pos;
        }
    }
}

// Word fill on the device, used instead of hipMemsetAsync where a caller may
// capture the stream into a graph (the allocation-free entry point).
__global__ void __launch_bounds__(256) fill_u32_kernel(unsigned* p, int64_t n, unsigned v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}
__global__ void __launch_bounds__(256) fill_f64_kernel(double* p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// Device scratch of odemi_solve_batch_noalloc: one array per GPU, held until
// the process ends.  Host code never passes `scratch` itself to a kernel (on
// the host it names a stand-in, not device memory); it uses the address that
// hipGetSymbolAddress returned for the device, kept in g_scratch.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Look up this device's scratch address once (the device must be current).
int scratch_setup(int device, std::string& err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) { err = "device index out of range for the scratch table"; return -1; }
    if (g_scratch[device]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return -1;
    }
    g_scratch[device] = ptr;
    return 0;
}

void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL) {
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ULL; }
    return h;
}

std::string cache_dir() {
    const char* e = getenv("ODEMI_CACHE_DIR");
    if (e && *e) return e;
    const char* h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/odemi";
}

void mkdirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

}  // namespace

struct odemi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string err;
    std::mutex mu;
    // grow-only device workspace
    void* ws = nullptr;
    size_t ws_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_order = nullptr;   // ordering against the legacy default stream (no timing)
    std::string scratch_err;   // why the allocation-free entry point is unavailable, if it is
};

struct odemi_rhs {
    odemi_ctx* ctx = nullptr;
    hipModule_t mod = nullptr;
    int n = 0, np = 0, flags = 0;
    int features = 0;   // lane hooks 1 isoutofdomain, 2 discrete, 4 continuous callbacks; workgroup callbacks 8, 16
    hipFunction_t tsit5_persist_f64 = nullptr, tsit5_persist_mixed = nullptr;
    hipFunction_t tsit5_persist_f64_sv = nullptr, tsit5_persist_mixed_sv = nullptr;
    hipFunction_t tsit5_plain_init = nullptr, tsit5_plain_step = nullptr;
    hipFunction_t rodas_persist_f64 = nullptr, rodas_persist_mixed = nullptr;
    hipFunction_t rodas_plain_init = nullptr, rodas_plain_step = nullptr;
    hipFunction_t rodas_wg_f64 = nullptr, rodas_wg_mixed = nullptr;
    hipFunction_t rodas_wg_f64_sv = nullptr, rodas_wg_mixed_sv = nullptr;
    hipFunction_t serve_init[2][2] = {}, serve_tick[2][2] = {};  // [alg][fp64, mixed]
    hipFunction_t serve_tick_sv[2][2] = {};                      // the same, requests with save times
    hipFunction_t tsit5_prep[2] = {}, rodas_prep = nullptr;      // start-up kernels [fp64, mixed]
    hipFunction_t rec_unpack = nullptr;                          // queue-order results -> SoA outputs
    hipFunction_t wg_serve_init[2] = {}, wg_serve_tick[2] = {};  // workgroup serving [fp64, mixed]
};

#define HIPCHK(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

extern "C" int odemi_abi_version(void) { return 4; }

extern "C" odemi_ctx* odemi_create(int device) {
    odemi_ctx* c = new odemi_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { c->err = "hipSetDevice failed"; return c; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) c->err = "stream create failed";
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    (void)hipEventCreateWithFlags(&c->ev_order, hipEventDisableTiming);
    // the allocation-free entry point's scratch address on this device
    (void)scratch_setup(device, c->scratch_err);
    return c;
}

extern "C" void odemi_destroy(odemi_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->ws) (void)hipFree(c->ws);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_order) (void)hipEventDestroy(c->ev_order);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* odemi_last_error(const odemi_ctx* c) { return c ? c->err.c_str() : "null ctx"; }

extern "C" int odemi_device_info(odemi_ctx* c, int* cu, size_t* fr, size_t* tot) {
    (void)hipSetDevice(c->device);
    if (cu) *cu = c->cu_count;
    size_t f = 0, t = 0;
    HIPCHK(c, hipMemGetInfo(&f, &t));
    if (fr) *fr = f;
    if (tot) *tot = t;
    return 0;
}

extern "C" void odemi_default_args(odemi_solve_args* a, int alg) {
    memset(a, 0, sizeof(*a));
    a->alg = alg;
    a->engine = ODEMI_ENGINE_PERSISTENT;
    a->precision = ODEMI_FP64;
    a->reltol = 1e-3;
    a->abstol = 1e-6;
    a->dtmax = 0.0;
    a->dtmin = 0.0;
    a->dt = 0.0;
    a->maxiters = 1000000;
    a->save_start = 1;
    a->save_end = 1;
    a->save_everystep = 0;
    a->save_end_user = -1;
    a->refine_sweeps = 1;
}

static std::string build_source(const char* problem_src, int n, int np, int flags) {
    std::ostringstream s;
    s << "#define ODEMI_N " << n << "\n#define ODEMI_NP " << (np > 0 ? np : 1) << "\n";
    s << "#define ODEMI_HAS_TGRAD " << ((flags & ODEMI_HAS_USER_TGRAD) ? 1 : 0) << "\n";
    // minimum waves per SIMD for the persistent Tsit5 lane kernels (occupancy vs
    // spills): $ODEMI_MIN_WAVES, else chosen after the problem source below
    const char* mw = getenv("ODEMI_MIN_WAVES");
    if (mw && *mw) s << "#define ODEMI_MIN_WAVES " << atoi(mw) << "\n";
    // the mixed lane Rodas5P persistent kernel: 3 waves per SIMD (168 VGPRs, a
    // few spills) measured 20 % faster than 2 (198 VGPRs) and than 4 (128 VGPRs,
    // heavy spills); the FP64 kernel gains nothing from a cap
    const char* rw = getenv("ODEMI_RODAS_MIN_WAVES");
    s << "#define ODEMI_RODAS_MIN_WAVES " << ((rw && *rw) ? atoi(rw) : 3) << "\n";
    // queue slots a wave takes per atomic (>= 64)
    const char* qc = getenv("ODEMI_QUEUE_CHUNK");
    if (qc && *qc && atoi(qc) >= 64) s << "#define ODEMI_QUEUE_CHUNK " << atoi(qc) << "\n";
    // idle lanes a wave gathers before refilling them (1..64)
    const char* rm = getenv("ODEMI_REFILL_MIN");
    if (rm && *rm && atoi(rm) >= 1 && atoi(rm) <= 64) s << "#define ODEMI_REFILL_MIN " << atoi(rm) << "\n";
    if (flags & ODEMI_COMPILE_WG) s << "#define ODEMI_WITH_WG 1\n";
    // Gauss-Jordan form of the mixed workgroup kernels: 1 blocked with MFMA
    // rank-16 updates (default), 0 element-wise rank-1 updates
    const char* gm = getenv("ODEMI_GJ_MFMA");
    if (gm && *gm) s << "#define ODEMI_GJ_MFMA " << (atoi(gm) ? 1 : 0) << "\n";
    s << problem_src << "\n";
    // 4 waves per SIMD (128 VGPRs); with continuous callbacks the event search
    // spills at that cap, and 3 measured 8 % faster (1M bouncing balls, FP64
    // engine: 1.91 -> 1.76 ms)
    s << "#ifndef ODEMI_MIN_WAVES\n#ifdef ODEMI_NCCB\n#define ODEMI_MIN_WAVES 3\n#else\n"
         "#define ODEMI_MIN_WAVES 4\n#endif\n#endif\n";
    s << kDeviceSource << "\n";
    // layout check of the kernel argument struct against this file's mirror
    s << "extern \"C\" __device__ const unsigned odemi_params_bytes = sizeof(odemi::Params);\n";
    // per-step hooks of the problem source, which only the lane kernels run
    // (and the workgroup kernels' callbacks: 8 discrete, 16 continuous)
    s << "extern \"C\" __device__ const int odemi_features = 0\n#ifdef ODEMI_HAS_ISOUT\n | 1\n#endif\n"
         "#ifdef ODEMI_NDCB\n | 2\n#endif\n#ifdef ODEMI_NCCB\n | 4\n#endif\n"
         "#ifdef ODEMI_NDCB_WG\n | 8\n#endif\n#ifdef ODEMI_NCCB_WG\n | 16\n#endif\n;\n";
    if (flags & ODEMI_COMPILE_TSIT5) {
        // persistent kernels: final-state-only (no save path) and saving variants
        const char* v[4][3] = {{"odemi_tsit5_persist_f64", "false", "double"},
                               {"odemi_tsit5_persist_f64_sv", "true", "double"},
                               {"odemi_tsit5_persist_mixed", "false", "float"},
                               {"odemi_tsit5_persist_mixed_sv", "true", "float"}};
        for (auto& k : v)
            s << "extern \"C\" __global__ void __launch_bounds__(256, ODEMI_MIN_WAVES) " << k[0]
              << "(odemi::Params P) { odemi::tsit5_persistent<" << k[1] << ", double, " << k[2]
              << ", ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_plain_init(odemi::Params P)"
             " { odemi::tsit5_plain_init<ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_prep_f64(odemi::Params P)"
             " { odemi::tsit5_prep<double, double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_prep_mixed(odemi::Params P)"
             " { odemi::tsit5_prep<double, float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rec_unpack(odemi::Params P)"
             " { odemi::rec_unpack<ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_init_f64(odemi::Params P)"
             " { odemi::tsit5_serve_init<double, double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_init_mixed(odemi::Params P)"
             " { odemi::tsit5_serve_init<double, float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_tick_f64(odemi::Params P)"
             " { odemi::tsit5_serve_tick<false, double, double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_tick_mixed(odemi::Params P)"
             " { odemi::tsit5_serve_tick<false, double, float, ODEMI_N>(P); }\n";
        // servers whose requests may carry save times
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_tick_f64_sv(odemi::Params P)"
             " { odemi::tsit5_serve_tick<true, double, double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_serve_tick_mixed_sv(odemi::Params P)"
             " { odemi::tsit5_serve_tick<true, double, float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_tsit5_plain_step(odemi::Params P)"
             " { odemi::tsit5_plain_step<ODEMI_N>(P); }\n";
    }
    if (flags & ODEMI_COMPILE_RODAS5P) {
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_persist_f64(odemi::Params P)"
             " { odemi::rodas_persistent<double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256, ODEMI_RODAS_MIN_WAVES) odemi_rodas_persist_mixed(odemi::Params P)"
             " { odemi::rodas_persistent<float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_plain_init(odemi::Params P)"
             " { odemi::rodas_plain_init<ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_prep(odemi::Params P)"
             " { odemi::rodas_prep<ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_serve_init_f64(odemi::Params P)"
             " { odemi::rodas_serve_init<double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_serve_init_mixed(odemi::Params P)"
             " { odemi::rodas_serve_init<float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_serve_tick_f64(odemi::Params P)"
             " { odemi::rodas_serve_tick<double, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_serve_tick_mixed(odemi::Params P)"
             " { odemi::rodas_serve_tick<float, ODEMI_N>(P); }\n";
        s << "extern \"C\" __global__ void __launch_bounds__(256) odemi_rodas_plain_step(odemi::Params P)"
             " { odemi::rodas_plain_step<ODEMI_N>(P); }\n";
    }
    if (flags & ODEMI_COMPILE_WG) {
        // the mixed variant uses the register-tiled Gauss-Jordan solver when n <= 128
        const char* v[4][4] = {{"odemi_rodas_wg_f64", "double", "false", "false"},
                               {"odemi_rodas_wg_f64_sv", "double", "true", "false"},
                               {"odemi_rodas_wg_mixed", "float", "false", "(ODEMI_GJ != 0)"},
                               {"odemi_rodas_wg_mixed_sv", "float", "true", "(ODEMI_GJ != 0)"}};
        // register cap for the mixed variant: minimum waves per SIMD (more
        // workgroups per CU against a few spills; 3 measured best at n = 128)
        const char* ww = getenv("ODEMI_WG_MIN_WAVES");
        const int wgw = (ww && *ww) ? atoi(ww) : 3;
        for (auto& k : v) {
            s << "extern \"C\" __global__ void __launch_bounds__(ODEMI_WG_THREADS) ";
            if (wgw > 0 && k[1][0] == 'f') s << "__attribute__((amdgpu_waves_per_eu(" << wgw << "))) ";
            s << k[0] << "(odemi::Params P) { odemi::rodas_wg_impl<" << k[1] << ", " << k[2] << ", " << k[3]
              << ">(P); }\n";
        }
        const char* sv[4][4] = {{"odemi_rodas_wg_serve_init_f64", "rodas_wg_serve_init", "double", "false"},
                                {"odemi_rodas_wg_serve_tick_f64", "rodas_wg_serve_tick", "double", "false"},
                                {"odemi_rodas_wg_serve_init_mixed", "rodas_wg_serve_init", "float", "(ODEMI_GJ != 0)"},
                                {"odemi_rodas_wg_serve_tick_mixed", "rodas_wg_serve_tick", "float", "(ODEMI_GJ != 0)"}};
        for (auto& k : sv) {
            s << "extern \"C\" __global__ void __launch_bounds__(ODEMI_WG_THREADS) ";
            if (wgw > 0 && k[2][0] == 'f') s << "__attribute__((amdgpu_waves_per_eu(" << wgw << "))) ";
            s << k[0] << "(odemi::Params P) { odemi::" << k[1] << "<" << k[2] << ", " << k[3] << ">(P); }\n";
        }
    }
    return s.str();
}

static const char* kRtcOpts[] = {
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};

static bool rtc_compile(const std::string& src, std::vector<char>& code, std::string& log) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "odemi_problem.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    int nopt = (int)(sizeof(kRtcOpts) / sizeof(kRtcOpts[0]));
    hiprtcResult r = hiprtcCompileProgram(prog, nopt, kRtcOpts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (ls > 1) {
        std::vector<char> l(ls + 1, 0);
        hiprtcGetProgramLog(prog, l.data());
        log = l.data();
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

// The file of the hipRTC library this process compiles with (two ROCm builds
// report the same hiprtcVersion, so the cache key also names the library).
static std::string rtc_library_path() {
    Dl_info info;
    if (dladdr((void*)&hiprtcVersion, &info) == 0 || !info.dli_fname) return "unknown";
    char buf[PATH_MAX];
    return realpath(info.dli_fname, buf) ? std::string(buf) : std::string(info.dli_fname);
}

extern "C" int odemi_rtc_version(int* major, int* minor) {
    int a = 0, b = 0;
    if (hiprtcVersion(&a, &b) != HIPRTC_SUCCESS) return -1;
    if (major) *major = a;
    if (minor) *minor = b;
    return 0;
}

// Compile (or fetch from the disk cache) the code object for a problem.
// Exposed so that a build step can pre-populate the cache without a GPU.
extern "C" int odemi_rtc_codeobject(const char* problem_src, int n, int np, int flags, char* path_out,
                                    int path_cap, double* seconds, char* log_out, int log_cap) {
    std::string src = build_source(problem_src, n, np, flags);
    // the key covers everything that determines the code: the full source
    // (device headers included), the options (target arch included) and the
    // hipRTC major and minor version
    std::string key = src;
    for (const char* o : kRtcOpts) key += o;
    int vmaj = 0, vmin = 0;
    if (hiprtcVersion(&vmaj, &vmin) != HIPRTC_SUCCESS) vmaj = vmin = -1;
    key += "|hiprtc " + std::to_string(vmaj) + "." + std::to_string(vmin) + " " + rtc_library_path();
    char hex[32];
    snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)fnv1a(key));
    std::string dir = cache_dir();
    std::string path = dir + "/odemi_" + hex + ".co";
    auto t0 = std::chrono::steady_clock::now();
    struct stat st;
    if (stat(path.c_str(), &st) != 0) {
        std::vector<char> code;
        std::string log;
        if (!rtc_compile(src, code, log)) {
            if (log_out && log_cap > 0) { strncpy(log_out, log.c_str(), log_cap - 1); log_out[log_cap - 1] = 0; }
            return -1;
        }
        mkdirs(dir);
        std::string tmp = path + ".tmp" + std::to_string(getpid());
        std::ofstream f(tmp, std::ios::binary);
        f.write(code.data(), (std::streamsize)code.size());
        f.close();
        rename(tmp.c_str(), path.c_str());
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (path_out && path_cap > 0) { strncpy(path_out, path.c_str(), path_cap - 1); path_out[path_cap - 1] = 0; }
    return 0;
}

extern "C" odemi_rhs* odemi_compile_rhs(odemi_ctx* ctx, const char* problem_src, int n, int np, int flags,
                                        double* compile_seconds) {
    if (!ctx) return nullptr;
    if (n < 1) { ctx->err = "n must be >= 1"; return nullptr; }
    if (!(flags & (ODEMI_COMPILE_TSIT5 | ODEMI_COMPILE_RODAS5P | ODEMI_COMPILE_WG))) flags |= ODEMI_COMPILE_TSIT5;
    char path[4096];
    std::vector<char> log(65536, 0);   // per call: contexts may compile on different threads
    double secs = 0;
    if (odemi_rtc_codeobject(problem_src, n, np, flags, path, sizeof(path), &secs, log.data(), (int)log.size()) != 0) {
        ctx->err = std::string("hipRTC compile failed:\n") + log.data();
        return nullptr;
    }
    if (compile_seconds) *compile_seconds = secs;
    std::ifstream f(path, std::ios::binary);
    std::vector<char> code((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    (void)hipSetDevice(ctx->device);
    odemi_rhs* r = new odemi_rhs();
    r->ctx = ctx;
    r->n = n;
    r->np = np;
    r->flags = flags;
    hipError_t e = hipModuleLoadData(&r->mod, code.data());
    if (e != hipSuccess) {
        ctx->err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        delete r;
        return nullptr;
    }
    {
        // the kernels take Params by value: its two declarations must agree
        hipDeviceptr_t gp = nullptr;
        size_t gsz = 0;
        unsigned dev_bytes = 0;
        if (hipModuleGetGlobal(&gp, &gsz, r->mod, "odemi_params_bytes") != hipSuccess || gsz != sizeof(unsigned) ||
            hipMemcpyDtoH(&dev_bytes, gp, sizeof(unsigned)) != hipSuccess || dev_bytes != sizeof(Params)) {
            ctx->err = "kernel Params layout differs from the runtime's (device " + std::to_string(dev_bytes) +
                       " bytes, host " + std::to_string(sizeof(Params)) + ")";
            (void)hipModuleUnload(r->mod);
            delete r;
            return nullptr;
        }
        if (hipModuleGetGlobal(&gp, &gsz, r->mod, "odemi_features") != hipSuccess || gsz != sizeof(int) ||
            hipMemcpyDtoH(&r->features, gp, sizeof(int)) != hipSuccess) {
            ctx->err = "module lacks the odemi_features word";
            (void)hipModuleUnload(r->mod);
            delete r;
            return nullptr;
        }
    }
    auto get = [&](const char* name, hipFunction_t* fn) {
        if (hipModuleGetFunction(fn, r->mod, name) != hipSuccess) *fn = nullptr;
    };
    if (flags & ODEMI_COMPILE_TSIT5) {
        get("odemi_tsit5_persist_f64", &r->tsit5_persist_f64);
        get("odemi_tsit5_persist_mixed", &r->tsit5_persist_mixed);
        get("odemi_tsit5_persist_f64_sv", &r->tsit5_persist_f64_sv);
        get("odemi_tsit5_persist_mixed_sv", &r->tsit5_persist_mixed_sv);
        get("odemi_tsit5_serve_init_f64", &r->serve_init[ODEMI_TSIT5][0]);
        get("odemi_tsit5_serve_init_mixed", &r->serve_init[ODEMI_TSIT5][1]);
        get("odemi_tsit5_serve_tick_f64", &r->serve_tick[ODEMI_TSIT5][0]);
        get("odemi_tsit5_serve_tick_mixed", &r->serve_tick[ODEMI_TSIT5][1]);
        get("odemi_tsit5_serve_tick_f64_sv", &r->serve_tick_sv[ODEMI_TSIT5][0]);
        get("odemi_tsit5_serve_tick_mixed_sv", &r->serve_tick_sv[ODEMI_TSIT5][1]);
        get("odemi_tsit5_plain_init", &r->tsit5_plain_init);
        get("odemi_tsit5_prep_f64", &r->tsit5_prep[0]);
        get("odemi_tsit5_prep_mixed", &r->tsit5_prep[1]);
        get("odemi_rec_unpack", &r->rec_unpack);
        get("odemi_tsit5_plain_step", &r->tsit5_plain_step);
    }
    if (flags & ODEMI_COMPILE_RODAS5P) {
        get("odemi_rodas_persist_f64", &r->rodas_persist_f64);
        get("odemi_rodas_persist_mixed", &r->rodas_persist_mixed);
        get("odemi_rodas_plain_init", &r->rodas_plain_init);
        get("odemi_rodas_prep", &r->rodas_prep);
        get("odemi_rodas_plain_step", &r->rodas_plain_step);
        get("odemi_rodas_serve_init_f64", &r->serve_init[ODEMI_RODAS5P][0]);
        get("odemi_rodas_serve_init_mixed", &r->serve_init[ODEMI_RODAS5P][1]);
        get("odemi_rodas_serve_tick_f64", &r->serve_tick[ODEMI_RODAS5P][0]);
        get("odemi_rodas_serve_tick_mixed", &r->serve_tick[ODEMI_RODAS5P][1]);
        // the Rosenbrock lane step saves on a runtime test: one kernel for both
        r->serve_tick_sv[ODEMI_RODAS5P][0] = r->serve_tick[ODEMI_RODAS5P][0];
        r->serve_tick_sv[ODEMI_RODAS5P][1] = r->serve_tick[ODEMI_RODAS5P][1];
    }
    if (flags & ODEMI_COMPILE_WG) {
        get("odemi_rodas_wg_f64", &r->rodas_wg_f64);
        get("odemi_rodas_wg_mixed", &r->rodas_wg_mixed);
        get("odemi_rodas_wg_f64_sv", &r->rodas_wg_f64_sv);
        get("odemi_rodas_wg_mixed_sv", &r->rodas_wg_mixed_sv);
        get("odemi_rodas_wg_serve_init_f64", &r->wg_serve_init[0]);
        get("odemi_rodas_wg_serve_init_mixed", &r->wg_serve_init[1]);
        get("odemi_rodas_wg_serve_tick_f64", &r->wg_serve_tick[0]);
        get("odemi_rodas_wg_serve_tick_mixed", &r->wg_serve_tick[1]);
    }
    return r;
}

extern "C" void odemi_rhs_destroy(odemi_rhs* r) {
    if (!r) return;
    if (r->mod) {
        (void)hipSetDevice(r->ctx->device);
        (void)hipModuleUnload(r->mod);
    }
    delete r;
}

namespace {

// Bump allocator over the ctx workspace (256-byte aligned slices).
struct Arena {
    char* base;
    size_t cap, off = 0;
    void* take(size_t bytes) {
        size_t o = (off + 255) & ~size_t(255);
        if (o + bytes > cap) return nullptr;
        off = o + bytes;
        return base + o;
    }
};

int ensure_ws(odemi_ctx* c, size_t bytes) {
    if (c->ws_bytes >= bytes) return 0;
    if (c->ws) { (void)hipFree(c->ws); c->ws = nullptr; c->ws_bytes = 0; }
    size_t b = std::max(bytes, (size_t)1 << 20);
    HIPCHK(c, hipMalloc(&c->ws, b));
    c->ws_bytes = b;
    return 0;
}

// Row stride (and side) of each workgroup's FP32 Jacobian block in the mixed
// workgroup kernels: 16 ceil(nmax / 16), as ODEMI_GJ_LD in kernels/odemi_gj.h.
size_t gj_ld(int nmax) { return 16 * (((size_t)nmax + 15) / 16); }

// Dynamic LDS of the workgroup kernels; must match odemi::wg_layout in kernels/odemi_wg.h.
size_t wg_shmem_bytes(int nmax, bool mixed, bool saves, bool gj) {
    const size_t ld = (size_t)nmax + 1, fb = mixed ? 4 : 8;
    const size_t np16 = 16 * (((size_t)nmax + 15) / 16);
    size_t shmem = gj ? 0 : (((size_t)nmax * ld * fb + 15) & ~(size_t)15);
    shmem += (size_t)(13 + (saves ? 3 : 0)) * nmax * 8 + 32 * 8;  // vectors + parameters (<= 32)
    shmem += (size_t)(nmax + 32) * 4 + (size_t)nmax * 8;          // fsc + pf, piv + perm
    shmem = (shmem + 15) & ~(size_t)15;
    if (gj) shmem += (51 * np16 + 16) * 4;                        // Gauss-Jordan buffers + zbuf
    shmem = (shmem + 15) & ~(size_t)15;
    shmem += 32 * 8 + 32 * 4;
    return shmem;
}

int launch(odemi_ctx* c, hipFunction_t fn, int grid, int block, Params& P, hipStream_t s, size_t shmem = 0) {
    void* args[] = {&P};
    HIPCHK(c, hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)shmem, s, args, nullptr));
    return 0;
}

// Zero `words` 32-bit words: hipMemsetAsync, or (allocation-free path, whose
// stream a caller may capture) the fill kernel.
int clear_words(odemi_ctx* c, void* p, size_t words, hipStream_t s, bool kernel) {
    if (words == 0) return 0;
    if (!kernel) {
        HIPCHK(c, hipMemsetAsync(p, 0, words * 4, s));
        return 0;
    }
    const int64_t g = std::min<int64_t>(((int64_t)words + 255) / 256, 1024);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)g), dim3(256), 0, s, (unsigned*)p, (int64_t)words, 0u);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// One ensemble solve.  noalloc: device pointers and the persistent engines
// only, every scratch buffer carved from this device's `scratch` array,
// nothing allocated, asynchronous on the stream (no synchronisation, no
// kernel_ms).
int solve_impl(odemi_ctx* ctx, odemi_rhs* rhs, odemi_solve_args* a, bool noalloc) {
    if (!ctx || !rhs || !a) return -1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    const int64_t N = a->N;
    const int nmax = a->nmax;
    const int np = a->np;
    a->kernel_ms = 0.0;
    a->launches = 0;
    if (noalloc) {
        if (!a->device_ptrs) { ctx->err = "the allocation-free entry point takes device pointers only"; return -2; }
        if (a->engine == ODEMI_ENGINE_PLAIN) {
            ctx->err = "the allocation-free entry point runs the persistent engines only";
            return -2;
        }
        if (!scratch_ptr(ctx->device)) {
            ctx->err = "no scratch on this device: " + ctx->scratch_err;
            return -1;
        }
    }
    if (N <= 0) return 0;
    if (nmax != rhs->n) { ctx->err = "args.nmax does not match the compiled problem dimension"; return -2; }
    if (np > std::max(rhs->np, 1)) { ctx->err = "args.np exceeds the compiled parameter count"; return -2; }
    if ((rhs->flags & ODEMI_COMPILE_WG) && np > 32) { ctx->err = "workgroup kernels take at most 32 parameters"; return -2; }
    if (a->max_saves < 0 || (a->max_saves > 0 && (!a->t_out || !a->u_out))) { ctx->err = "save buffers missing"; return -2; }
    if (!a->retcode) { ctx->err = "retcode output is required"; return -2; }
    if (a->nsaveat < 0 || (a->nsaveat > 0 && !a->saveat)) { ctx->err = "saveat buffer missing"; return -2; }
    for (int i = 1; i < a->nsaveat; i++)
        if (!a->device_ptrs && a->saveat[i] < a->saveat[i - 1]) { ctx->err = "saveat must be sorted"; return -2; }
    if (a->ntstops < 0 || (a->ntstops > 0 && !a->tstops)) { ctx->err = "tstops buffer missing"; return -2; }
    for (int i = 1; i < a->ntstops; i++)
        if (!a->device_ptrs && a->tstops[i] < a->tstops[i - 1]) { ctx->err = "tstops must be sorted"; return -2; }
    const bool rodas = (a->alg == ODEMI_RODAS5P);
    const bool plain = (a->engine == ODEMI_ENGINE_PLAIN);
    const bool mixed = (a->precision == ODEMI_MIXED);
    const bool use_wg = !plain && rodas && (rhs->flags & ODEMI_COMPILE_WG);
    // plain engine of a workgroup problem: the serving pool's workgroup kernels
    // with one step attempt per launch, state in HBM between launches (FP64 LU)
    const bool plain_wg = plain && rodas && (rhs->flags & ODEMI_COMPILE_WG);
    hipFunction_t fn_main = nullptr, fn_init = nullptr, fn_prep = nullptr;
    if (!rodas) {
        if (plain) { fn_init = rhs->tsit5_plain_init; fn_main = rhs->tsit5_plain_step; }
        else if (a->nsaveat > 0 || a->save_everystep)
            fn_main = mixed ? rhs->tsit5_persist_mixed_sv : rhs->tsit5_persist_f64_sv;
        else fn_main = mixed ? rhs->tsit5_persist_mixed : rhs->tsit5_persist_f64;
        if (!plain) fn_prep = rhs->tsit5_prep[mixed ? 1 : 0];
    } else if (plain_wg) {
        fn_init = rhs->wg_serve_init[0]; fn_main = rhs->wg_serve_tick[0];
        if (a->nsaveat > 0 || a->save_everystep) {
            ctx->err = "the plain engine of uneven-size batches saves the start and end states only (no saveat, save_everystep)";
            return -3;
        }
    } else if (use_wg) {
        if (a->nsaveat > 0 || a->save_everystep) fn_main = mixed ? rhs->rodas_wg_mixed_sv : rhs->rodas_wg_f64_sv;
        else fn_main = mixed ? rhs->rodas_wg_mixed : rhs->rodas_wg_f64;
    } else {
        if (plain) { fn_init = rhs->rodas_plain_init; fn_main = rhs->rodas_plain_step; }
        else fn_main = mixed ? rhs->rodas_persist_mixed : rhs->rodas_persist_f64;
        if (!plain) fn_prep = rhs->rodas_prep;
    }
    if (plain && mixed) { ctx->err = "the plain engine is FP64 only"; return -3; }
    // queue-order records (final-state-only Tsit5 persistent engine, no save
    // before the end point, no lane callbacks; not in the allocation-free form;
    // $ODEMI_QUEUE_RECORDS=0 reads and writes the SoA arrays in place instead)
    const char* qr = getenv("ODEMI_QUEUE_RECORDS");
    const bool use_rec = !(qr && atoi(qr) == 0) && !rodas && !plain && !noalloc && fn_prep && rhs->rec_unpack &&
                         a->sort_by_cost && N > 0 &&
                         a->nsaveat == 0 && !a->save_everystep && !a->save_start && !(rhs->features & 6) &&
                         nmax == rhs->n;
    const int rec_np = std::max(rhs->np, 1);
    if (!fn_main || (plain && !fn_init)) { ctx->err = "requested kernel was not compiled for this problem"; return -3; }
    if (a->nvec && !use_wg && !plain_wg) { ctx->err = "uneven batches (nvec) need the workgroup Rodas5P kernels"; return -3; }
    if ((use_wg || plain_wg) && (rhs->features & kLaneHooks)) {
        ctx->err = "the lane kernels' isoutofdomain and callback forms need a lane problem (compile without ODEMI_COMPILE_WG)";
        return -3;
    }

    hipStream_t s = a->stream ? (hipStream_t)a->stream : ctx->stream;
    const bool dev = a->device_ptrs != 0;
    // Device arrays with no stream given: they may come from work on the legacy
    // default stream, which the ctx stream (non-blocking) does not wait for.
    const bool order_null = dev && !a->stream;
    if (order_null) {
        HIPCHK(ctx, hipEventRecord(ctx->ev_order, nullptr));
        HIPCHK(ctx, hipStreamWaitEvent(s, ctx->ev_order, 0));
    }
    const size_t N_ = (size_t)N;
    const size_t ms_ = (size_t)std::max(a->max_saves, 0);
    // plain-engine vector rows: Tsit5 u, uprev, k1, k7; Rodas5P lane u, uprev, dT, J;
    // workgroup u, uprev, dT
    const size_t vrows = plain_wg ? 3 : rodas ? (size_t)(3 + nmax) : 4;
    // ---- launch geometry of the persistent kernels ----
    const int block = a->block > 0 ? a->block : 256;
    const int threads = use_wg ? 256 : block;   // ODEMI_WG_THREADS for the workgroup kernels
    const bool use_gj = use_wg && mixed && nmax <= 128;   // register-tiled Gauss-Jordan variant
    size_t shmem = 0;
    int64_t grid = 0;
    if (!plain) {
        if (use_wg) {
            shmem = wg_shmem_bytes(nmax, mixed, a->nsaveat > 0 || a->save_everystep || (rhs->features & 16), use_gj);
            if (shmem > 160 * 1024) { ctx->err = "state dimension too large for the workgroup kernel's LDS"; return -3; }
        }
        int nb = 0;
        HIPCHK(ctx, hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn_main, threads, shmem));
        // Resident workgroups per CU: args.blocks_per_cu (the autotuner's
        // choice), else $ODEMI_BLOCKS_PER_CU, else for the lane kernels at most
        // 4 (16 waves): they are VALU-bound at that occupancy, and fewer lanes
        // give each lane more trajectories, so the longest-first tail is
        // shorter (Lorenz 1M mixed: 5 per CU 2.44 ms, 4 per CU 2.37 ms, 3 per
        // CU 2.42 ms).  Never above what occupancy allows.
        const char* bpc = getenv("ODEMI_BLOCKS_PER_CU");
        if (a->blocks_per_cu > 0) nb = std::min(nb, a->blocks_per_cu);
        else if (bpc && atoi(bpc) > 0) nb = std::min(nb, atoi(bpc));
        else if (!use_wg) nb = std::min(nb, 4);
        if (nb < 1) nb = 1;
        grid = (int64_t)nb * std::max(ctx->cu_count, 1);
        const int64_t per_block = use_wg ? 1 : threads;
        grid = std::min<int64_t>(grid, (N + per_block - 1) / per_block);
    } else if (plain_wg) {
        // the tick kernels' layout keeps the dense rows when continuous callbacks search events
        shmem = wg_shmem_bytes(nmax, false, (rhs->features & 16) != 0, false);
        if (shmem > 160 * 1024) { ctx->err = "state dimension too large for the workgroup kernel's LDS"; return -3; }
        int nb = 0;
        HIPCHK(ctx, hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn_main, 256, shmem));
        grid = std::min<int64_t>((int64_t)std::max(nb, 1) * std::max(ctx->cu_count, 1), N);
    }
    // ---- workspace layout ----
    size_t need = 0;
    auto add = [&](size_t b) { need += ((b + 255) & ~size_t(255)); };
    add(64);                                   // counters
    if (a->sort_by_cost && !plain) {
        if (N > 0x7fffffffLL) { ctx->err = "N too large for the work-queue order"; return -2; }
        add(N_ * 4); add(kCostBins * 4);
    }
    if (plain) { add(16 * N_ * 8); add(vrows * nmax * N_ * 8); add(kIntRows * N_ * 4); }
    if (plain_wg) { add(N_ * 4); if (!a->nvec) add(N_ * 4); }   // slot -> request (all 0); nvec = nmax
    if (use_gj) add((size_t)grid * gj_ld(nmax) * gj_ld(nmax) * 4);
    if (fn_prep) { add(N_ * 8); if (!rodas) add(N_ * nmax * 8); }
    if (use_rec) { add(N_ * 4); add(N_ * rec_in_words(nmax, rec_np) * 8); add(N_ * rec_out_words(nmax) * 8); }
    if (!dev) {
        add(N_ * nmax * 8); add(N_ * std::max(np, 1) * 8); add(N_ * 8); add(N_ * 8);
        add(std::max(a->nsaveat, 1) * 8);
        add(std::max(a->ntstops, 1) * 8);
        add(N_ * std::max<size_t>(ms_, 1) * 8); add(N_ * std::max<size_t>(ms_, 1) * nmax * 8);
        add(N_ * 4); add(N_ * nmax * 8); add(N_ * 8); add(N_ * 4); add(N_ * kStatRows * 4);
        if (a->nvec) add(N_ * 4);
    }
    Arena ar{nullptr, 0};
    if (noalloc) {
        if (need > kScratchBytes) {
            ctx->err = "this solve needs " + std::to_string(need) + " bytes of scratch; the allocation-free entry point has " +
                       std::to_string(kScratchBytes);
            return -2;
        }
        ar = Arena{(char*)scratch_ptr(ctx->device), kScratchBytes};
    } else {
        if (ensure_ws(ctx, need)) return -1;
        ar = Arena{(char*)ctx->ws, ctx->ws_bytes};
    }
    Params P;
    memset(&P, 0, sizeof(P));
    P.N = N; P.nmax = nmax; P.np = np;
    P.reltol = a->reltol; P.abstol = a->abstol; P.dtmax = a->dtmax; P.dtmin = a->dtmin; P.dt_user = a->dt;
    P.maxiters = a->maxiters;
    P.nsaveat = a->nsaveat;
    P.ntstops = a->ntstops;
    P.save_start = a->save_start; P.save_end = a->save_end; P.save_everystep = a->save_everystep;
    P.save_end_user = a->save_end_user; P.max_saves = a->max_saves;
    P.refine_sweeps = a->refine_sweeps;
    P.counters = (unsigned long long*)ar.take(64);
    if (clear_words(ctx, P.counters, 16, s, noalloc)) return -1;
    if (dev) {
        P.u0 = a->u0; P.p = a->p; P.t0 = a->t0; P.tf = a->tf; P.saveat = a->saveat; P.nvec = a->nvec;
        P.tstops = a->tstops;
        P.t_out = a->t_out; P.u_out = a->u_out; P.nsaved = a->nsaved; P.u_final = a->u_final;
        P.t_final = a->t_final; P.retcode = a->retcode; P.stats = a->stats;
    } else {
        double* u0 = (double*)ar.take(N_ * nmax * 8);
        double* pp = (double*)ar.take(N_ * std::max(np, 1) * 8);
        double* t0 = (double*)ar.take(N_ * 8);
        double* tf = (double*)ar.take(N_ * 8);
        double* sa = (double*)ar.take(std::max(a->nsaveat, 1) * 8);
        HIPCHK(ctx, hipMemcpyAsync(u0, a->u0, N_ * nmax * 8, hipMemcpyHostToDevice, s));
        if (np > 0) HIPCHK(ctx, hipMemcpyAsync(pp, a->p, N_ * np * 8, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(t0, a->t0, N_ * 8, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(tf, a->tf, N_ * 8, hipMemcpyHostToDevice, s));
        if (a->nsaveat > 0) HIPCHK(ctx, hipMemcpyAsync(sa, a->saveat, a->nsaveat * 8, hipMemcpyHostToDevice, s));
        double* ts = (double*)ar.take(std::max(a->ntstops, 1) * 8);
        if (a->ntstops > 0) HIPCHK(ctx, hipMemcpyAsync(ts, a->tstops, a->ntstops * 8, hipMemcpyHostToDevice, s));
        P.u0 = u0; P.p = pp; P.t0 = t0; P.tf = tf; P.saveat = sa; P.tstops = ts;
        P.t_out = (double*)ar.take(N_ * std::max<size_t>(ms_, 1) * 8);
        P.u_out = (double*)ar.take(N_ * std::max<size_t>(ms_, 1) * nmax * 8);
        P.nsaved = (int*)ar.take(N_ * 4);
        P.u_final = (double*)ar.take(N_ * nmax * 8);
        P.t_final = (double*)ar.take(N_ * 8);
        P.retcode = (int*)ar.take(N_ * 4);
        P.stats = (int*)ar.take(N_ * kStatRows * 4);
        if (a->nvec) {
            int* nv = (int*)ar.take(N_ * 4);
            HIPCHK(ctx, hipMemcpyAsync(nv, a->nvec, N_ * 4, hipMemcpyHostToDevice, s));
            P.nvec = nv;
        }
    }
    if (a->sort_by_cost && !plain && N > 0) {
        int* ord = (int*)ar.take(N_ * 4);
        unsigned* bins = (unsigned*)ar.take(kCostBins * 4);
        const int g = (int)((N + kCostTile - 1) / kCostTile);
        if (clear_words(ctx, bins, kCostBins, s, noalloc)) return -1;
        hipLaunchKernelGGL(cost_hist_kernel, dim3(g), dim3(256), 0, s, P.t0, P.tf, P.nvec, N, bins);
        hipLaunchKernelGGL(cost_scan_kernel, dim3(1), dim3(256), 0, s, bins);
        int* inv = use_rec ? (int*)ar.take(N_ * 4) : nullptr;
        hipLaunchKernelGGL(cost_scatter_kernel, dim3(g), dim3(256), 0, s, P.t0, P.tf, P.nvec, N, bins, ord, inv);
        HIPCHK(ctx, hipGetLastError());
        P.order = ord;
        P.inv = inv;
    }
    if (plain) {
        P.st_scalar = (double*)ar.take(16 * N_ * 8);
        P.st_vec = (double*)ar.take(vrows * nmax * N_ * 8);
        P.st_int = (int*)ar.take(kIntRows * N_ * 4);
        HIPCHK(ctx, hipMemsetAsync(P.st_int, 0, kIntRows * N_ * 4, s));
    }
    if (plain_wg) {
        // every slot reports to request-table entry 0, whose done counter is counters[1]
        int* sreq = (int*)ar.take(N_ * 4);
        if (clear_words(ctx, sreq, N_, s, false)) return -1;
        P.slot_req = sreq;
        P.req_done = P.counters + 1;
        if (!P.nvec) {
            int* nv = (int*)ar.take(N_ * 4);
            hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 1024)), dim3(256), 0, s,
                               (unsigned*)nv, (int64_t)N, (unsigned)nmax);
            HIPCHK(ctx, hipGetLastError());
            P.nvec = nv;
        }
        P.slot0 = 0;
        P.nslots = N;
        P.tick_attempts = 1;
        // the workgroup tick kernels take each slot's tolerances from scalar rows 10 and 11
        const unsigned g = (unsigned)std::min<int64_t>((N + 255) / 256, 1024);
        hipLaunchKernelGGL(fill_f64_kernel, dim3(g), dim3(256), 0, s, P.st_scalar + 10 * N_, (int64_t)N, P.reltol);
        hipLaunchKernelGGL(fill_f64_kernel, dim3(g), dim3(256), 0, s, P.st_scalar + 11 * N_, (int64_t)N, P.abstol);
        HIPCHK(ctx, hipGetLastError());
    }
    if (use_gj) P.wg_jbuf = (float*)ar.take((size_t)grid * gj_ld(nmax) * gj_ld(nmax) * 4);
    if (use_rec) {
        P.rec_in = (double*)ar.take(N_ * rec_in_words(nmax, rec_np) * 8);
        P.rec_out = (double*)ar.take(N_ * rec_out_words(nmax) * 8);
        if (!P.inv || !P.rec_in || !P.rec_out) { ctx->err = "workspace too small for the queue-order records"; return -1; }
    }
    if (fn_prep) {
        P.pre_dt = (double*)ar.take(N_ * 8);
        if (!rodas) P.pre_k1 = (double*)ar.take(N_ * nmax * 8);
    }
    if (P.nsaved && clear_words(ctx, P.nsaved, N_, s, noalloc)) return -1;

    int64_t launches = 0;
    // kernel_ms: the plain engine's launches, or the persistent kernel alone
    // (the prep kernel runs before the first event)
    if (plain) HIPCHK(ctx, hipEventRecord(ctx->ev0, s));
    if (plain) {
        // The outer loop: batches of `kBatch` step launches, the host reading
        // the live (lane engine) or finished (workgroup engine) count after
        // each.  A batch is captured once per solve into a hipGraph and
        // replayed ($ODEMI_PLAIN_GRAPH=0: launched one by one).  Counters are
        // cleared by the fill kernel, so the graph holds kernel nodes only.
        const int lgrid = (int)((N + block - 1) / block);
        if (plain_wg) {
            // init: one workgroup per trajectory (grid-strided)
            if (launch(ctx, fn_init, (int)grid, 256, P, s, shmem)) return -1;
        } else {
            if (launch(ctx, fn_init, lgrid, block, P, s)) return -1;
        }
        launches++;
        constexpr int kBatch = 32;
        // plain_wg: each launch takes one attempt of every live trajectory,
        // handed out by counters[0]; lane engine: counters[1] counts the lanes
        // still live after the batch's last launch
        auto enqueue_batch = [&]() -> int {
            for (int k = 0; k < kBatch; k++) {
                if (plain_wg) {
                    if (clear_words(ctx, P.counters, 2, s, true)) return -1;
                    if (launch(ctx, fn_main, (int)grid, 256, P, s, shmem)) return -1;
                } else {
                    if (k == kBatch - 1 && clear_words(ctx, P.counters + 1, 2, s, true)) return -1;
                    if (launch(ctx, fn_main, lgrid, block, P, s)) return -1;
                }
            }
            return 0;
        };
        const char* pg = getenv("ODEMI_PLAIN_GRAPH");
        const bool use_graph = !(pg && atoi(pg) == 0);
        hipGraphExec_t exec = nullptr;
        if (use_graph) {
            hipGraph_t g = nullptr;
            HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_batch();
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (!rc) ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
                return -1;
            }
            const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) { ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei); return -1; }
        }
        unsigned long long cnt = 0;
        int rc = 0;
        for (;;) {
            if (use_graph) {
                const hipError_t e = hipGraphLaunch(exec, s);
                if (e != hipSuccess) { ctx->err = std::string("hipGraphLaunch: ") + hipGetErrorString(e); rc = -1; break; }
            } else if (enqueue_batch()) {
                rc = -1;
                break;
            }
            launches += kBatch;
            hipError_t e = hipMemcpyAsync(&cnt, P.counters + 1, 8, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) { ctx->err = std::string("plain engine count: ") + hipGetErrorString(e); rc = -1; break; }
            if (plain_wg ? (cnt >= (unsigned long long)N) : (cnt == 0)) break;
            if (launches > (int64_t)a->maxiters * 4 + 1024) { ctx->err = "plain engine did not terminate"; rc = -4; break; }
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (rc) return rc;
    } else {
        if (fn_prep) {
            if (launch(ctx, fn_prep, (int)((N + 255) / 256), 256, P, s)) return -1;
            launches++;
        }
        if (!noalloc) HIPCHK(ctx, hipEventRecord(ctx->ev0, s));
        if (launch(ctx, fn_main, (int)grid, threads, P, s, shmem)) return -1;
        launches++;
    }
    if (noalloc) {
        if (order_null) {
            HIPCHK(ctx, hipEventRecord(ctx->ev_order, s));
            HIPCHK(ctx, hipStreamWaitEvent(nullptr, ctx->ev_order, 0));
        }
        a->launches = launches;
        return 0;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, s));
    if (use_rec) {
        if (launch(ctx, rhs->rec_unpack, (int)((N + 255) / 256), 256, P, s)) return -1;
        launches++;
    }
    if (!dev) {
        if (ms_ > 0) {
            HIPCHK(ctx, hipMemcpyAsync(a->t_out, P.t_out, N_ * ms_ * 8, hipMemcpyDeviceToHost, s));
            HIPCHK(ctx, hipMemcpyAsync(a->u_out, P.u_out, N_ * ms_ * nmax * 8, hipMemcpyDeviceToHost, s));
        }
        if (a->nsaved) HIPCHK(ctx, hipMemcpyAsync(a->nsaved, P.nsaved, N_ * 4, hipMemcpyDeviceToHost, s));
        if (a->u_final) HIPCHK(ctx, hipMemcpyAsync(a->u_final, P.u_final, N_ * nmax * 8, hipMemcpyDeviceToHost, s));
        if (a->t_final) HIPCHK(ctx, hipMemcpyAsync(a->t_final, P.t_final, N_ * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(a->retcode, P.retcode, N_ * 4, hipMemcpyDeviceToHost, s));
        if (a->stats) HIPCHK(ctx, hipMemcpyAsync(a->stats, P.stats, N_ * kStatRows * 4, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    a->kernel_ms = ms;
    a->launches = launches;
    return 0;
}

}  // namespace

extern "C" int odemi_solve_batch(odemi_ctx* ctx, odemi_rhs* rhs, odemi_solve_args* a) {
    return solve_impl(ctx, rhs, a, false);
}

extern "C" int odemi_solve_batch_noalloc(odemi_ctx* ctx, odemi_rhs* rhs, odemi_solve_args* a) {
    return solve_impl(ctx, rhs, a, true);
}

// ===========================================================================
// Serving: continuous batching over a device-resident slot pool
// ===========================================================================
namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0, N = 0;
    std::vector<double> u0, p, t0, tf;   // host copies of the inputs (SoA, N columns)
    std::vector<double> tol;             // reltol [N] then abstol [N] (the request's own or the server's)
    std::vector<int> nvec;               // per-trajectory n (workgroup servers)
    std::vector<int> sreq;               // slot -> request-table index for the admitted range
    double* u_final = nullptr;           // caller's output buffers
    int32_t* retcode = nullptr;
    int32_t* stats = nullptr;
    std::vector<double> sa, ts;          // the request's save times and stopping times (sorted)
    int ropt[kReqOpts] = {};      // its request-table entry: nsaveat, ntstops, save flags
    double* t_out = nullptr;             // caller's buffers for the saved points (or null)
    double* u_out = nullptr;
    int32_t* nsaved = nullptr;
    int64_t slot0 = -1;
    int rt = -1;                         // request-table index on the device
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;                       // 0 pending, 1 active, 2 done, <0 error
};

// First-fit allocator of contiguous slot ranges.
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;    // start -> length
    int64_t high = 0;                    // high-water mark (grid extent)
    explicit RangeAlloc(int64_t n = 0) { if (n > 0) free_[0] = n; }
    int64_t take(int64_t n) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                int64_t s = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[s + n] = len - n;
                high = std::max(high, s + n);
                return s;
            }
        }
        return -1;
    }
    void give(int64_t s, int64_t n) {
        auto it = free_.emplace(s, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); it = pv; }
        }
        // shrink the high-water mark when the tail is free
        auto last = std::prev(free_.end());
        if (last->first + last->second >= high) high = std::min(high, last->first);
    }
};
}  // namespace

struct odemi_server {
    odemi_ctx* ctx = nullptr;
    odemi_rhs* rhs = nullptr;
    odemi_server_opts o;
    int64_t S = 0;                       // slots
    int R = 4096;                        // request table entries
    int rows = 4;
    bool wg = false, gj = false;         // workgroup kernels (uneven n), Gauss-Jordan variant
    size_t shmem = 0;
    int wg_grid = 0;                     // resident workgroups of the tick kernel
    hipStream_t stream = nullptr;
    void* mem = nullptr;
    Params P;
    std::vector<unsigned long long> done_h;
    RangeAlloc alloc;
    std::vector<int> rt_free;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::map<int64_t, Request*> all;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::string err;
    int64_t ticks = 0, admitted = 0, finished = 0;
};

static void server_loop(odemi_server* S);

extern "C" void odemi_server_default_opts(odemi_server_opts* o, int alg) {
    memset(o, 0, sizeof(*o));
    o->alg = alg;
    o->precision = ODEMI_FP64;
    o->reltol = 1e-3;
    o->abstol = 1e-6;
    o->maxiters = 1000000;
    o->tick_attempts = 64;
    o->max_slots = 1 << 20;
    o->refine_sweeps = 1;
}

extern "C" odemi_server* odemi_server_create(odemi_ctx* ctx, odemi_rhs* rhs, const odemi_server_opts* opts) {
    if (!ctx || !rhs || !opts) return nullptr;
    const int pi = opts->precision == ODEMI_MIXED ? 1 : 0;
    const bool wg = (rhs->flags & ODEMI_COMPILE_WG) && opts->alg == ODEMI_RODAS5P;
    if (opts->alg < 0 || opts->alg > 1 || !(wg ? rhs->wg_serve_tick[pi] : rhs->serve_tick[opts->alg][pi])) {
        ctx->err = "the compiled problem has no serving kernels for this algorithm";
        return nullptr;
    }
    if (wg && (rhs->features & kLaneHooks)) {
        ctx->err = "the lane kernels' isoutofdomain and callback forms need a lane problem";
        return nullptr;
    }
    if (opts->max_saves < 0 || opts->max_tstops < 0) { ctx->err = "max_saves and max_tstops must be >= 0"; return nullptr; }
    if (wg && opts->max_saves > 0) {
        ctx->err = "save times need a lane (fixed-n) server";
        return nullptr;
    }
    if (opts->max_saves > 0 && opts->alg == ODEMI_TSIT5 && !rhs->serve_tick_sv[ODEMI_TSIT5][pi]) {
        ctx->err = "the compiled problem has no saving serve kernels";
        return nullptr;
    }
    (void)hipSetDevice(ctx->device);
    odemi_server* S = new odemi_server();
    S->ctx = ctx;
    S->rhs = rhs;
    S->o = *opts;
    const int n = rhs->n, np = std::max(rhs->np, 1);
    S->wg = wg;
    S->gj = wg && pi == 1 && n <= 128;
    S->rows = wg ? 3 : (opts->alg == ODEMI_RODAS5P) ? 3 + n : 4;
    if (wg) {
        // resident workgroups of the tick kernel (the Gauss-Jordan variant needs a
        // Jacobian block per workgroup)
        S->shmem = wg_shmem_bytes(n, pi == 1, (rhs->features & 16) != 0, S->gj);   // dense rows for events
        if (S->shmem > 160 * 1024) { ctx->err = "state dimension too large for the workgroup kernel's LDS"; delete S; return nullptr; }
        int nb = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rhs->wg_serve_tick[pi], 256, S->shmem) != hipSuccess)
            nb = 1;
        S->wg_grid = std::max(nb, 1) * std::max(ctx->cu_count, 1);
    }
    // bytes per slot: inputs, state, outputs
    const int msv = opts->max_saves, mts = opts->max_tstops;
    const size_t per_slot = (size_t)(n + np + 2) * 8 + 16 * 8 + (size_t)S->rows * n * 8 + kIntRows * 4 +
                            (size_t)n * 8 + 8 + 4 + 4 + kStatRows * 4 + 4 + (wg ? 4 : 0) +
                            (size_t)msv * (n + 1) * 8;   // saved points
    const size_t fixed = S->gj ? (size_t)S->wg_grid * gj_ld(n) * gj_ld(n) * 4 : 0;
    int64_t slots = opts->max_slots > 0 ? opts->max_slots : (1 << 20);
    if (opts->hbm_budget_bytes > 0)
        slots = std::min<int64_t>(slots, opts->hbm_budget_bytes > fixed
                                             ? (int64_t)((opts->hbm_budget_bytes - fixed) / per_slot) : 0);
    if (slots < 1) { ctx->err = "HBM budget too small for one slot"; delete S; return nullptr; }
    S->S = slots;
    const size_t Sz = (size_t)slots;
    size_t bytes = 0;
    auto al = [&](size_t b) { size_t o = bytes; bytes += (b + 255) & ~size_t(255); return o; };
    size_t o_u0 = al(Sz * n * 8), o_p = al(Sz * np * 8), o_t0 = al(Sz * 8), o_tf = al(Sz * 8);
    size_t o_sc = al(16 * Sz * 8), o_vec = al((size_t)S->rows * n * Sz * 8), o_int = al(kIntRows * Sz * 4);
    size_t o_uf = al(Sz * n * 8), o_tfin = al(Sz * 8), o_rc = al(Sz * 4), o_st = al(kStatRows * Sz * 4), o_ns = al(Sz * 4);
    size_t o_sreq = al(Sz * 4), o_done = al((size_t)S->R * 8), o_cnt = al(64);
    size_t o_nvec = wg ? al(Sz * 4) : 0, o_jb = S->gj ? al((size_t)S->wg_grid * gj_ld(n) * gj_ld(n) * 4) : 0;
    // saved points of every slot, and the request table's save / stopping times and flags
    const bool reqtab = msv > 0 || mts > 0;
    size_t o_tout = msv > 0 ? al((size_t)msv * Sz * 8) : 0, o_uout = msv > 0 ? al((size_t)msv * n * Sz * 8) : 0;
    size_t o_rsa = reqtab ? al((size_t)S->R * std::max(msv, 1) * 8) : 0;
    size_t o_rts = reqtab ? al((size_t)S->R * std::max(mts, 1) * 8) : 0;
    size_t o_ropt = reqtab ? al((size_t)S->R * kReqOpts * 4) : 0;
    if (hipMalloc(&S->mem, bytes) != hipSuccess) { ctx->err = "serving pool allocation failed"; delete S; return nullptr; }
    char* b = (char*)S->mem;
    (void)hipMemset(b + o_int, 0, kIntRows * Sz * 4);
    (void)hipMemset(b + o_done, 0, (size_t)S->R * 8);
    (void)hipMemset(b + o_cnt, 0, 64);
    Params& P = S->P;
    memset(&P, 0, sizeof(P));
    P.N = slots; P.nmax = n; P.np = rhs->np;
    P.u0 = (double*)(b + o_u0); P.p = (double*)(b + o_p); P.t0 = (double*)(b + o_t0); P.tf = (double*)(b + o_tf);
    P.reltol = opts->reltol; P.abstol = opts->abstol; P.dtmax = opts->dtmax; P.dtmin = opts->dtmin; P.dt_user = 0.0;
    P.maxiters = opts->maxiters;
    P.save_start = 0; P.save_end = 0; P.save_everystep = 0; P.save_end_user = -1; P.max_saves = 0;
    P.st_scalar = (double*)(b + o_sc); P.st_vec = (double*)(b + o_vec); P.st_int = (int*)(b + o_int);
    P.u_final = (double*)(b + o_uf); P.t_final = (double*)(b + o_tfin); P.retcode = (int*)(b + o_rc);
    P.stats = (int*)(b + o_st); P.nsaved = (int*)(b + o_ns);
    P.slot_req = (int*)(b + o_sreq); P.req_done = (unsigned long long*)(b + o_done);
    P.counters = (unsigned long long*)(b + o_cnt);
    if (wg) P.nvec = (const int*)(b + o_nvec);
    if (S->gj) P.wg_jbuf = (float*)(b + o_jb);
    if (msv > 0) {
        P.max_saves = msv;
        P.t_out = (double*)(b + o_tout);
        P.u_out = (double*)(b + o_uout);
    }
    if (reqtab) {
        (void)hipMemset(b + o_ropt, 0, (size_t)S->R * kReqOpts * 4);
        P.req_saveat = (const double*)(b + o_rsa);
        P.req_tstops = (const double*)(b + o_rts);
        P.req_opts = (const int*)(b + o_ropt);
        P.sa_cap = std::max(msv, 1);
        P.ts_cap = std::max(mts, 1);
    }
    P.refine_sweeps = opts->refine_sweeps;
    P.tick_attempts = opts->tick_attempts > 0 ? opts->tick_attempts : 64;
    S->done_h.assign(S->R, 0);
    S->alloc = RangeAlloc(slots);
    for (int i = S->R - 1; i >= 0; i--) S->rt_free.push_back(i);
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
        ctx->err = "stream create failed"; (void)hipFree(S->mem); delete S; return nullptr;
    }
    S->worker = std::thread(server_loop, S);
    return S;
}

extern "C" void odemi_server_destroy(odemi_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    {
        // requests that never finished end as failed (-4); no other call may be
        // in progress on this server from here on
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& kv : S->all)
            if (kv.second->state == 0 || kv.second->state == 1) kv.second->state = -4;
        S->cv_done.notify_all();
    }
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->stream);
    (void)hipStreamDestroy(S->stream);
    (void)hipFree(S->mem);
    for (auto& kv : S->all) delete kv.second;
    delete S;
}

extern "C" int odemi_request_release(odemi_server* S, int64_t id) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return -1;
    if (it->second->state == 0 || it->second->state == 1) return -2;   // still queued or running
    delete it->second;
    S->all.erase(it);
    return 0;
}

extern "C" int64_t odemi_submit(odemi_server* S, int64_t N, const int32_t* nvec, const double* u0, const double* p,
                                const double* t0, const double* tf, double* u_final, int32_t* retcode,
                                int32_t* stats) {
    return odemi_submit_opts(S, N, nvec, u0, p, t0, tf, nullptr, u_final, retcode, stats);
}

extern "C" int64_t odemi_submit_opts(odemi_server* S, int64_t N, const int32_t* nvec, const double* u0,
                                     const double* p, const double* t0, const double* tf,
                                     const odemi_request_opts* ro, double* u_final, int32_t* retcode,
                                     int32_t* stats) {
    if (!S || N <= 0 || !u0 || !t0 || !tf || !retcode) return -1;
    if (S->rhs->np > 0 && !p) { S->err = "the problem has parameters but p is NULL"; return -1; }
    if (N > S->S) { S->err = "request larger than the slot pool"; return -2; }
    const int n = S->rhs->n, np = S->rhs->np;
    Request* r = new Request();
    r->N = N;
    r->u0.assign(u0, u0 + (size_t)n * N);
    if (np > 0 && p) r->p.assign(p, p + (size_t)np * N); else r->p.assign((size_t)std::max(np, 1) * N, 0.0);
    r->t0.assign(t0, t0 + N);
    r->tf.assign(tf, tf + N);
    const double rt = (ro && ro->reltol > 0.0) ? ro->reltol : S->o.reltol;
    const double at = (ro && ro->abstol > 0.0) ? ro->abstol : S->o.abstol;
    r->tol.assign((size_t)2 * N, rt);
    std::fill(r->tol.begin() + N, r->tol.end(), at);
    if (S->wg) {
        r->nvec.assign(N, n);
        if (nvec) {
            for (int64_t k = 0; k < N; k++) {
                if (nvec[k] < 1 || nvec[k] > n) { S->err = "nvec entries must lie in [1, n]"; delete r; return -3; }
                r->nvec[k] = nvec[k];
            }
        }
    } else if (nvec) {
        for (int64_t k = 0; k < N; k++)
            if (nvec[k] != n) { S->err = "uneven nvec needs a workgroup (Rodas5P) server"; delete r; return -3; }
    }
    r->u_final = u_final; r->retcode = retcode; r->stats = stats;
    if (ro) {
        const bool saves = ro->t_out != nullptr;
        if (ro->nsaveat < 0 || ro->ntstops < 0 || (ro->nsaveat > 0 && !ro->saveat) || (ro->ntstops > 0 && !ro->tstops)) {
            S->err = "saveat / tstops: negative count or NULL array"; delete r; return -3;
        }
        if ((ro->nsaveat > 0 || saves) && S->wg) {
            S->err = "save times need a lane (fixed-n) server"; delete r; return -3;
        }
        if (saves) {
            if (!ro->u_out) { S->err = "t_out without u_out"; delete r; return -3; }
            const int64_t need = ro->nsaveat + (ro->save_start ? 1 : 0) + (ro->save_end ? 1 : 0);
            if (need > S->o.max_saves) {
                S->err = "the request asks for " + std::to_string(need) + " saved points; the server keeps " +
                         std::to_string(S->o.max_saves);
                delete r; return -3;
            }
            r->sa.assign(ro->saveat, ro->saveat + ro->nsaveat);
            std::sort(r->sa.begin(), r->sa.end());
            r->ropt[0] = (int)ro->nsaveat;
            r->ropt[2] = ro->save_start ? 1 : 0;
            r->ropt[3] = ro->save_end ? 1 : 0;
            r->t_out = ro->t_out; r->u_out = ro->u_out; r->nsaved = ro->nsaved;
        } else if (ro->nsaveat > 0) {
            S->err = "saveat without t_out / u_out"; delete r; return -3;
        }
        r->ropt[4] = ro->save_end_user;
        if (ro->ntstops > 0) {
            if (ro->ntstops > S->o.max_tstops) {
                S->err = "the request has " + std::to_string(ro->ntstops) + " stopping times; the server keeps " +
                         std::to_string(S->o.max_tstops);
                delete r; return -3;
            }
            r->ts.assign(ro->tstops, ro->tstops + ro->ntstops);
            std::sort(r->ts.begin(), r->ts.end());
            r->ropt[1] = (int)ro->ntstops;
        }
    } else {
        r->ropt[4] = -1;
    }
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) { S->err = "the server has stopped"; delete r; return -4; }
        r->id = S->next_id++;
        S->all[r->id] = r;
        S->pending.push_back(r);
    }
    S->cv.notify_all();
    return r->id;
}

extern "C" int odemi_poll(odemi_server* S, int64_t id) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return -1;
    return it->second->state == 2 ? 1 : (it->second->state < 0 ? it->second->state : 0);
}

extern "C" int odemi_wait(odemi_server* S, int64_t id, double timeout_s) {
    if (!S) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return -1;
    Request* r = it->second;
    auto pred = [&] { return r->state == 2 || r->state < 0; };
    if (timeout_s > 0) {
        if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return 0;
    } else {
        S->cv_done.wait(lk, pred);
    }
    return r->state == 2 ? 1 : r->state;
}

extern "C" int odemi_request_times(odemi_server* S, int64_t id, double* t_submit, double* t_admit, double* t_done) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return -1;
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_admit) *t_admit = it->second->t_admit;
    if (t_done) *t_done = it->second->t_done;
    return 0;
}

extern "C" int odemi_server_info(odemi_server* S, int64_t* slots, int64_t* ticks, int64_t* pending, int64_t* active,
                                 int64_t* finished) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (slots) *slots = S->S;
    if (ticks) *ticks = S->ticks;
    if (pending) *pending = (int64_t)S->pending.size();
    if (active) *active = (int64_t)S->active.size();
    if (finished) *finished = S->finished;
    return 0;
}

extern "C" const char* odemi_server_last_error(odemi_server* S) { return S ? S->err.c_str() : "null server"; }

static int server_fail(odemi_server* S, const char* what, hipError_t e) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = std::string(what) + ": " + hipGetErrorString(e);
    for (auto* r : S->active) r->state = -3;
    for (auto* r : S->pending) r->state = -3;
    S->active.clear();
    S->pending.clear();
    S->stop = true;
    S->cv_done.notify_all();
    return -1;
}

#define SCHK(call)                                            \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) { server_fail(S, #call, e_); return; } \
    } while (0)

static void server_loop(odemi_server* S) {
    (void)hipSetDevice(S->ctx->device);
    const int n = S->rhs->n, np = std::max(S->rhs->np, 1);
    const int pi = S->o.precision == ODEMI_MIXED ? 1 : 0;
    hipFunction_t f_init = S->wg ? S->rhs->wg_serve_init[pi] : S->rhs->serve_init[S->o.alg][pi];
    hipFunction_t f_tick = S->wg ? S->rhs->wg_serve_tick[pi]
                                 : (S->o.max_saves > 0 ? S->rhs->serve_tick_sv[S->o.alg][pi]
                                                       : S->rhs->serve_tick[S->o.alg][pi]);
    const size_t Sz = (size_t)S->S;
    for (;;) {
        std::vector<Request*> admit;
        int64_t high = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->active.empty(); });
            if (S->stop && S->active.empty() && S->pending.empty()) break;
            if (S->stop) break;
            // admission: FIFO while a contiguous slot range and a request-table entry are free
            while (!S->pending.empty() && !S->rt_free.empty()) {
                Request* r = S->pending.front();
                int64_t s0 = S->alloc.take(r->N);
                if (s0 < 0) break;
                S->pending.pop_front();
                r->slot0 = s0;
                r->rt = S->rt_free.back();
                S->rt_free.pop_back();
                r->state = 1;
                r->t_admit = now_s();
                S->active.push_back(r);
                admit.push_back(r);
                S->admitted++;
            }
            high = S->alloc.high;
        }
        for (Request* r : admit) {
            const size_t N = (size_t)r->N, s0 = (size_t)r->slot0;
            SCHK(hipMemcpy2DAsync((double*)S->P.u0 + s0, Sz * 8, r->u0.data(), N * 8, N * 8, n, hipMemcpyHostToDevice, S->stream));
            SCHK(hipMemcpy2DAsync((double*)S->P.p + s0, Sz * 8, r->p.data(), N * 8, N * 8, np, hipMemcpyHostToDevice, S->stream));
            SCHK(hipMemcpyAsync((double*)S->P.t0 + s0, r->t0.data(), N * 8, hipMemcpyHostToDevice, S->stream));
            SCHK(hipMemcpyAsync((double*)S->P.tf + s0, r->tf.data(), N * 8, hipMemcpyHostToDevice, S->stream));
            // the slots' tolerances: scalar state rows 10 (reltol) and 11 (abstol)
            SCHK(hipMemcpyAsync(S->P.st_scalar + 10 * Sz + s0, r->tol.data(), N * 8, hipMemcpyHostToDevice, S->stream));
            SCHK(hipMemcpyAsync(S->P.st_scalar + 11 * Sz + s0, r->tol.data() + N, N * 8, hipMemcpyHostToDevice,
                                S->stream));
            r->sreq.assign(N, r->rt);
            SCHK(hipMemcpyAsync((int*)S->P.slot_req + s0, r->sreq.data(), N * 4, hipMemcpyHostToDevice, S->stream));
            if (S->P.req_opts) {   // the request's save / stopping times and flags
                if (!r->sa.empty())
                    SCHK(hipMemcpyAsync((double*)S->P.req_saveat + (size_t)r->rt * S->P.sa_cap, r->sa.data(),
                                        r->sa.size() * 8, hipMemcpyHostToDevice, S->stream));
                if (!r->ts.empty())
                    SCHK(hipMemcpyAsync((double*)S->P.req_tstops + (size_t)r->rt * S->P.ts_cap, r->ts.data(),
                                        r->ts.size() * 8, hipMemcpyHostToDevice, S->stream));
                SCHK(hipMemcpyAsync((int*)S->P.req_opts + (size_t)r->rt * kReqOpts, r->ropt,
                                    sizeof(r->ropt), hipMemcpyHostToDevice, S->stream));
            }
            SCHK(hipMemsetAsync(S->P.req_done + r->rt, 0, 8, S->stream));
            SCHK(hipMemsetAsync((int*)S->P.nsaved + s0, 0, N * 4, S->stream));
            if (S->wg) SCHK(hipMemcpyAsync((int*)S->P.nvec + s0, r->nvec.data(), N * 4, hipMemcpyHostToDevice, S->stream));
            Params Pi = S->P;
            Pi.slot0 = (int64_t)s0;
            Pi.nslots = (int64_t)N;
            void* args[] = {&Pi};
            if (S->wg)   // one workgroup per admitted slot
                SCHK(hipModuleLaunchKernel(f_init, (unsigned)std::min<size_t>(N, (size_t)S->wg_grid), 1, 1, 256, 1, 1,
                                           (unsigned)S->shmem, S->stream, args, nullptr));
            else
                SCHK(hipModuleLaunchKernel(f_init, (unsigned)((N + 255) / 256), 1, 1, 256, 1, 1, 0, S->stream, args, nullptr));
            // the host input vectors live in the Request until it finishes, so no wait here
        }
        if (high > 0 && S->wg) {
            // workgroups take slots of [0, high) from counters[0]
            SCHK(hipMemsetAsync(S->P.counters, 0, 8, S->stream));
            Params Pt = S->P;
            Pt.slot0 = 0;
            Pt.nslots = high;
            void* args[] = {&Pt};
            SCHK(hipModuleLaunchKernel(f_tick, (unsigned)std::min<int64_t>(high, S->wg_grid), 1, 1, 256, 1, 1,
                                       (unsigned)S->shmem, S->stream, args, nullptr));
        } else if (high > 0) {
            Params Pt = S->P;
            void* args[] = {&Pt};
            SCHK(hipModuleLaunchKernel(f_tick, (unsigned)((high + 255) / 256), 1, 1, 256, 1, 1, 0, S->stream, args, nullptr));
        }
        SCHK(hipMemcpyAsync(S->done_h.data(), S->P.req_done, (size_t)S->R * 8, hipMemcpyDeviceToHost, S->stream));
        SCHK(hipStreamSynchronize(S->stream));
        std::vector<Request*> fin;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->ticks++;
            for (size_t i = 0; i < S->active.size();) {
                Request* r = S->active[i];
                if ((int64_t)S->done_h[r->rt] >= r->N) {
                    fin.push_back(r);
                    S->active[i] = S->active.back();
                    S->active.pop_back();
                } else {
                    i++;
                }
            }
        }
        for (Request* r : fin) {
            const size_t N = (size_t)r->N, s0 = (size_t)r->slot0;
            if (r->u_final)
                SCHK(hipMemcpy2DAsync(r->u_final, N * 8, (double*)S->P.u_final + s0, Sz * 8, N * 8, n, hipMemcpyDeviceToHost, S->stream));
            SCHK(hipMemcpyAsync(r->retcode, (int*)S->P.retcode + s0, N * 4, hipMemcpyDeviceToHost, S->stream));
            if (r->stats)
                SCHK(hipMemcpy2DAsync(r->stats, N * 4, (int*)S->P.stats + s0, Sz * 4, N * 4, kStatRows, hipMemcpyDeviceToHost, S->stream));
            if (r->t_out) {
                const size_t ms = (size_t)S->o.max_saves;
                SCHK(hipMemcpy2DAsync(r->t_out, N * 8, (double*)S->P.t_out + s0, Sz * 8, N * 8, ms, hipMemcpyDeviceToHost, S->stream));
                SCHK(hipMemcpy2DAsync(r->u_out, N * 8, (double*)S->P.u_out + s0, Sz * 8, N * 8, ms * n, hipMemcpyDeviceToHost, S->stream));
                if (r->nsaved)
                    SCHK(hipMemcpyAsync(r->nsaved, (int*)S->P.nsaved + s0, N * 4, hipMemcpyDeviceToHost, S->stream));
            }
        }
        if (!fin.empty()) SCHK(hipStreamSynchronize(S->stream));
        if (!fin.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            const double t = now_s();
            for (Request* r : fin) {
                S->alloc.give(r->slot0, r->N);
                S->rt_free.push_back(r->rt);
                r->t_done = t;
                r->state = 2;
                std::vector<double>().swap(r->u0);
                std::vector<double>().swap(r->p);
                std::vector<double>().swap(r->t0);
                std::vector<double>().swap(r->tf);
                std::vector<int>().swap(r->sreq);
                std::vector<int>().swap(r->nvec);
                S->finished++;
            }
            S->cv_done.notify_all();
        }
    }
}
