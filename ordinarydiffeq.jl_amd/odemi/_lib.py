"""ctypes binding of the C ABI in include/odemi.h (libodemi.so, built in-tree).

The binding is deliberately thin: plain pointers and sizes, no torch types.
Loading fails loudly when the native library is missing; there is no CPU path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libodemi.so")
DEFAULT_CACHE = os.path.join(_HERE, "_cache")

TSIT5, RODAS5P = 0, 1
ENGINE_PLAIN, ENGINE_PERSISTENT = 0, 1
FP64, MIXED = 0, 1
# rows of the per-trajectory stats output: nf naccept nreject njacs nw nsolve ncondition
NSTATS = 7
COMPILE_TSIT5, COMPILE_RODAS5P, HAS_USER_TGRAD, COMPILE_WG = 1, 2, 4, 8

P = ctypes.POINTER
c_d, c_i, c_i64, c_vp = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
ABI_VERSION = 4   # odemi_abi_version() of the library these structures mirror


class SolveArgs(ctypes.Structure):
    """Mirror of odemi_solve_args (include/odemi.h).  Array fields are raw
    addresses (host or device, per device_ptrs)."""
    _fields_ = [
        ("alg", c_i), ("engine", c_i), ("precision", c_i), ("device_ptrs", c_i),
        ("N", c_i64), ("nmax", c_i), ("np", c_i),
        ("nvec", c_vp), ("u0", c_vp), ("p", c_vp), ("t0", c_vp), ("tf", c_vp),
        ("reltol", c_d), ("abstol", c_d), ("dtmax", c_d), ("dtmin", c_d), ("dt", c_d),
        ("maxiters", c_i64),
        ("saveat", c_vp), ("nsaveat", c_i),
        ("tstops", c_vp), ("ntstops", c_i),
        ("save_start", c_i), ("save_end", c_i), ("save_everystep", c_i), ("save_end_user", c_i),
        ("max_saves", c_i),
        ("t_out", c_vp), ("u_out", c_vp), ("nsaved", c_vp), ("u_final", c_vp), ("t_final", c_vp),
        ("retcode", c_vp), ("stats", c_vp),
        ("refine_sweeps", c_i), ("block", c_i), ("blocks_per_cu", c_i), ("sort_by_cost", c_i),
        ("stream", c_vp),
        ("kernel_ms", c_d), ("launches", c_i64),
    ]


class ServerOpts(ctypes.Structure):
    """Mirror of odemi_server_opts (include/odemi.h)."""
    _fields_ = [
        ("alg", c_i), ("precision", c_i),
        ("reltol", c_d), ("abstol", c_d), ("dtmax", c_d), ("dtmin", c_d),
        ("maxiters", c_i64), ("max_slots", c_i64), ("hbm_budget_bytes", ctypes.c_size_t),
        ("tick_attempts", c_i), ("refine_sweeps", c_i), ("max_saves", c_i), ("max_tstops", c_i),
    ]


class RequestOpts(ctypes.Structure):
    """Mirror of odemi_request_opts (include/odemi.h)."""
    _fields_ = [("reltol", c_d), ("abstol", c_d), ("saveat", c_vp), ("nsaveat", c_i64),
                ("save_start", c_i), ("save_end", c_i), ("save_end_user", c_i),
                ("t_out", c_vp), ("u_out", c_vp), ("nsaved", c_vp), ("tstops", c_vp), ("ntstops", c_i64)]


EXPORTS = ["odemi_create", "odemi_destroy", "odemi_last_error", "odemi_device_info", "odemi_abi_version",
           "odemi_compile_rhs", "odemi_rhs_destroy", "odemi_solve_batch", "odemi_default_args",
           "odemi_rtc_codeobject", "odemi_server_default_opts", "odemi_server_create", "odemi_server_destroy",
           "odemi_submit", "odemi_poll", "odemi_wait", "odemi_request_times", "odemi_server_info",
           "odemi_server_last_error", "odemi_request_release", "odemi_rtc_version", "odemi_submit_opts", "odemi_solve_batch_noalloc"]

_lib = None


def load():
    """Load libodemi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"odemi native library missing: {LIB_PATH} (run __graft_entry__.build())")
    os.environ.setdefault("ODEMI_CACHE_DIR", DEFAULT_CACHE)
    # One HIP runtime per process.  PyTorch ships its own ROCm build; loaded
    # after libodemi.so it brings a second HIP runtime into the process, and of
    # two runtimes only the one initialised first finds the GPU.  So torch, when
    # installed, is loaded first and libodemi.so binds to its runtime (the same
    # soname); without torch the system ROCm runtime is used.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    L.odemi_create.argtypes = [c_i]
    L.odemi_create.restype = c_vp
    L.odemi_destroy.argtypes = [c_vp]
    L.odemi_last_error.argtypes = [c_vp]
    L.odemi_last_error.restype = ctypes.c_char_p
    L.odemi_device_info.argtypes = [c_vp, P(c_i), P(ctypes.c_size_t), P(ctypes.c_size_t)]
    L.odemi_abi_version.restype = c_i
    if L.odemi_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI {L.odemi_abi_version()}, this binding expects {ABI_VERSION} "
                           "(rebuild with __graft_entry__.build())")
    L.odemi_compile_rhs.argtypes = [c_vp, ctypes.c_char_p, c_i, c_i, c_i, P(c_d)]
    L.odemi_compile_rhs.restype = c_vp
    L.odemi_rhs_destroy.argtypes = [c_vp]
    L.odemi_solve_batch.argtypes = [c_vp, c_vp, P(SolveArgs)]
    L.odemi_solve_batch.restype = c_i
    L.odemi_solve_batch_noalloc.argtypes = [c_vp, c_vp, P(SolveArgs)]
    L.odemi_solve_batch_noalloc.restype = c_i
    L.odemi_default_args.argtypes = [P(SolveArgs), c_i]
    L.odemi_rtc_codeobject.argtypes = [ctypes.c_char_p, c_i, c_i, c_i, ctypes.c_char_p, c_i, P(c_d),
                                       ctypes.c_char_p, c_i]
    L.odemi_rtc_codeobject.restype = c_i
    L.odemi_server_default_opts.argtypes = [P(ServerOpts), c_i]
    L.odemi_server_create.argtypes = [c_vp, c_vp, P(ServerOpts)]
    L.odemi_server_create.restype = c_vp
    L.odemi_server_destroy.argtypes = [c_vp]
    L.odemi_submit.argtypes = [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]
    L.odemi_submit.restype = c_i64
    L.odemi_submit_opts.argtypes = [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, P(RequestOpts), c_vp, c_vp, c_vp]
    L.odemi_submit_opts.restype = c_i64
    L.odemi_poll.argtypes = [c_vp, c_i64]
    L.odemi_poll.restype = c_i
    L.odemi_wait.argtypes = [c_vp, c_i64, c_d]
    L.odemi_wait.restype = c_i
    L.odemi_request_times.argtypes = [c_vp, c_i64, P(c_d), P(c_d), P(c_d)]
    L.odemi_request_times.restype = c_i
    L.odemi_server_info.argtypes = [c_vp] + [P(c_i64)] * 5
    L.odemi_server_info.restype = c_i
    L.odemi_request_release.argtypes = [c_vp, c_i64]
    L.odemi_request_release.restype = c_i
    L.odemi_server_last_error.argtypes = [c_vp]
    L.odemi_server_last_error.restype = ctypes.c_char_p
    L.odemi_rtc_version.argtypes = [P(c_i), P(c_i)]
    L.odemi_rtc_version.restype = c_i
    _lib = L
    return L


def precompile(src, n, np_, flags):
    """hipRTC-compile a problem into the on-disk code-object cache (no GPU needed)."""
    L = load()
    path = ctypes.create_string_buffer(4096)
    log = ctypes.create_string_buffer(65536)
    secs = c_d()
    rc = L.odemi_rtc_codeobject(src.encode(), n, np_, flags, path, 4096, ctypes.byref(secs), log, 65536)
    if rc != 0:
        raise RuntimeError("hipRTC compile failed:\n" + log.value.decode(errors="replace"))
    return path.value.decode(), secs.value


def rtc_version():
    """(major, minor) of the hipRTC that compiles the problems' code objects."""
    a, b = c_i(), c_i()
    if load().odemi_rtc_version(ctypes.byref(a), ctypes.byref(b)) != 0:
        raise RuntimeError("hipRTC did not report its version")
    return a.value, b.value


class Context:
    """One device context (odemi_ctx*): stream, workspace, compiled problems."""

    def __init__(self, device=0):
        self.L = load()
        self.h = self.L.odemi_create(device)
        if not self.h:
            raise RuntimeError("odemi_create failed")
        err = self.L.odemi_last_error(self.h)
        if err:
            raise RuntimeError("odemi_create: " + err.decode())
        self.device = device
        self._rhs = {}

    def error(self):
        return self.L.odemi_last_error(self.h).decode(errors="replace")

    def device_info(self):
        cu = c_i()
        fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
        if self.L.odemi_device_info(self.h, ctypes.byref(cu), ctypes.byref(fr), ctypes.byref(tot)) != 0:
            raise RuntimeError(self.error())
        return dict(cu_count=cu.value, free_bytes=fr.value, total_bytes=tot.value)

    def compile(self, src, n, np_, flags):
        key = (src, n, np_, flags)
        r = self._rhs.get(key)
        if r is None:
            if os.environ.get("ODEMI_LOG_CODEOBJECTS"):
                # name the cached code object this module loads (the file name is its key's hash)
                path, _ = precompile(src, n, np_, flags)
                print(f"[odemi] code object {os.path.basename(path)} (n={n}, flags={flags})", flush=True)
            secs = c_d()
            h = self.L.odemi_compile_rhs(self.h, src.encode(), n, np_, flags, ctypes.byref(secs))
            if not h:
                raise RuntimeError(self.error())
            r = (h, secs.value)
            self._rhs[key] = r
        return r

    def solve(self, rhs_handle, args, noalloc=False):
        """odemi_solve_batch, or with noalloc odemi_solve_batch_noalloc (device
        pointers, persistent engine, asynchronous on args.stream)."""
        fn = self.L.odemi_solve_batch_noalloc if noalloc else self.L.odemi_solve_batch
        rc = fn(self.h, rhs_handle, ctypes.byref(args))
        if rc != 0:
            raise RuntimeError(f"{fn.__name__} failed ({rc}): {self.error()}")
        return args

    def default_args(self, alg):
        a = SolveArgs()
        self.L.odemi_default_args(ctypes.byref(a), alg)
        return a

    def close(self):
        if self.h:
            for h, _ in self._rhs.values():
                self.L.odemi_rhs_destroy(h)
            self._rhs.clear()
            self.L.odemi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
