"""ctypes binding of libocl_amd.so (the C ABI in include/ocl_amd.h).

There is no fallback: if the library is missing or the GPU is absent, calls raise.  Device memory is owned by
DeviceArray (freed when collected); host arrays are numpy.
"""
from __future__ import annotations

import ctypes
import os
import re
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libocl_amd.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "ocl_amd.h")

OCL_FORM_PLAIN = 0
OCL_FORM_FUSED = 1
OCL_FORM_SPLIT = 2
OCL_E_INVALID, OCL_E_HIP, OCL_E_UNSUPPORTED, OCL_E_NOMEM = -1, -2, -3, -4
MAX_STAGES = 16
TRACE = 4096
RED = {"nrm2": 0, "dotc": 1, "sum": 2, "asum": 3, "maxabs": 4}


class OclError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ocl error {code}: {msg}")
        self.code = code


class NestaParams(ctypes.Structure):
    """NestaUp::LaunchParameters, same field order (include/OpenCLIPER/processes/nesta/NestaUp.hpp:67-83)."""
    _fields_ = [("lambda_i", ctypes.c_double), ("mu_f", ctypes.c_double), ("La", ctypes.c_double),
                ("maxIntIter", ctypes.c_int), ("tolVar", ctypes.c_double), ("verbose", ctypes.c_int),
                ("maxIter", ctypes.c_int), ("stoptest", ctypes.c_int), ("miniter", ctypes.c_int)]


def nesta_params(lambda_i=1e-2, mu_f=3e-4, La=1.0, maxIntIter=4, tolVar=3e-4, verbose=1, maxIter=100, stoptest=1,
                 miniter=7) -> NestaParams:
    """Defaults are the values tests/MRIRecon.cpp:105-114 of the reference passes to NestaUp."""
    return NestaParams(lambda_i, mu_f, La, maxIntIter, tolVar, verbose, maxIter, stoptest, miniter)


class NestaReport(ctypes.Structure):
    _fields_ = [("normU", ctypes.c_double), ("maxXref", ctypes.c_double), ("maxUXref", ctypes.c_double),
                ("lambda_", ctypes.c_double), ("muf", ctypes.c_double), ("mu0", ctypes.c_double),
                ("gamma", ctypes.c_double), ("gammat", ctypes.c_double), ("final_tolVar", ctypes.c_double),
                ("final_mu", ctypes.c_double), ("normest_iters", ctypes.c_int), ("stages", ctypes.c_int),
                ("iters", ctypes.c_int * MAX_STAGES), ("ntrace", ctypes.c_int),
                ("trace_f", ctypes.c_double * TRACE), ("trace_l2", ctypes.c_double * TRACE),
                ("trace_l1", ctypes.c_double * TRACE)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("trace") and k != "iters"}
        d["lambda"] = d.pop("lambda_")
        d["iters"] = list(self.iters)[: self.stages]
        n = self.ntrace
        d["trace_f"] = np.array(self.trace_f[:n])
        d["trace_l2"] = np.array(self.trace_l2[:n])
        d["trace_l1"] = np.array(self.trace_l1[:n])
        return d


class Slice(ctypes.Structure):
    _fields_ = [("y", ctypes.c_void_p), ("S", ctypes.c_void_p), ("M", ctypes.c_void_p), ("x_out", ctypes.c_void_p),
                ("F", ctypes.c_int), ("C", ctypes.c_int), ("H", ctypes.c_int), ("W", ctypes.c_int)]


class McArgs(ctypes.Structure):
    """ocl_mc_args: device pointers to the deformed positions xn [2][F][H][W] float32 and the original mesh
    x [2][W][H] int32 (both 1-based), and the 1-based bound box."""
    _fields_ = [("xn", ctypes.c_void_p), ("x", ctypes.c_void_p), ("bound_box", ctypes.c_uint * 4)]


GWR_TRACE = 1024


class GwrParams(ctypes.Structure):
    """ocl_gwr_params: GroupwiseRegistration::InitParameters, same order."""
    _fields_ = [("W", ctypes.c_float), ("flagW", ctypes.c_int), ("radius", ctypes.c_uint), ("E", ctypes.c_uint),
                ("Dp", ctypes.c_int * 2), ("nmax", ctypes.c_uint), ("et", ctypes.c_float), ("eh", ctypes.c_float),
                ("lambda_", ctypes.c_float * 4)]


def gwr_params(W=1.0, flagW=True, radius=40, E=3, Dp=(4, 4), nmax=100, et=0.01, eh=0.005,
               lam=(0.0, 0.005, 0.0, 0.5)) -> GwrParams:
    """Defaults are the values of the reference's tests/MRIRecon.cpp:122-130."""
    p = GwrParams(W, int(bool(flagW)), radius, E)
    p.Dp[0], p.Dp[1] = Dp
    p.nmax, p.et, p.eh = nmax, et, eh
    for i, v in enumerate(lam):
        p.lambda_[i] = v
    return p


class GwrInfo(ctypes.Structure):
    _fields_ = [("Cp", ctypes.c_int), ("support", ctypes.c_int), ("box", ctypes.c_int), ("areaX", ctypes.c_int),
                ("use_lambda", ctypes.c_int), ("bound_box", ctypes.c_uint * 4), ("T_elems", ctypes.c_long),
                ("xn_elems", ctypes.c_long), ("mesh", ctypes.c_void_p)]


class GwrReport(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("ntrace", ctypes.c_int), ("H0", ctypes.c_float), ("Wn", ctypes.c_float),
                ("oldmax", ctypes.c_float), ("H", ctypes.c_float * GWR_TRACE), ("Tdif", ctypes.c_float * GWR_TRACE),
                ("Hdif", ctypes.c_float * GWR_TRACE), ("Wn_trace", ctypes.c_float * GWR_TRACE)]

    def as_dict(self):
        n = self.ntrace
        return {"iters": self.iters, "H0": self.H0, "Wn": self.Wn, "oldmax": self.oldmax,
                "H": np.array(self.H[: n + 1], dtype=np.float64), "Tdif": np.array(self.Tdif[:n], dtype=np.float64),
                "Hdif": np.array(self.Hdif[:n], dtype=np.float64),
                "Wn_trace": np.array(self.Wn_trace[:n], dtype=np.float64)}


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_long), ("failed", ctypes.c_long), ("queued", ctypes.c_long),
                ("active", ctypes.c_long), ("used_bytes", ctypes.c_size_t), ("peak_used_bytes", ctypes.c_size_t),
                ("budget_bytes", ctypes.c_size_t), ("slice_iterations", ctypes.c_long), ("launches", ctypes.c_long),
                ("busy_s", ctypes.c_double), ("elapsed_s", ctypes.c_double), ("budget_waits", ctypes.c_long),
                ("cache_hits", ctypes.c_long), ("cache_misses", ctypes.c_long), ("admit_s", ctypes.c_double),
                ("retire_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SERVE_QUEUED, SERVE_RUNNING, SERVE_DONE = 0, 1, 2

_lib = None


def header_symbols():
    """Every function the C header declares."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void\*?|const char\*)\s+\**(ocl_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OclError(-100, f"{LIB_PATH} is not built; run python -m opencliper_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    P, I, LG, D, F32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "ocl_last_error": ([], ctypes.c_char_p),
        "ocl_ctx_create": ([I, PP], I), "ocl_ctx_destroy": ([P], I), "ocl_ctx_sync": ([P], I),
        "ocl_ctx_stream": ([P], P),
        "ocl_device_info": ([P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                             ctypes.POINTER(I)], I),
        "ocl_malloc": ([P, ctypes.c_size_t, PP], I), "ocl_free": ([P, P], I),
        "ocl_host_malloc": ([P, ctypes.c_size_t, PP], I), "ocl_host_free": ([P, P], I),
        "ocl_memcpy_h2d": ([P, P, P, ctypes.c_size_t], I), "ocl_memcpy_d2h": ([P, P, P, ctypes.c_size_t], I),
        "ocl_memcpy_d2d": ([P, P, P, ctypes.c_size_t], I), "ocl_memset_zero": ([P, P, ctypes.c_size_t], I),
        "ocl_ttv": ([P, P, P, I, LG, I], I), "ocl_vecnorm": ([P, P, P, LG, F32], I),
        "ocl_cep": ([P, P, P, P, I, I, LG, I, I], I), "ocl_apply_mask": ([P, P, P, I, I, LG], I),
        "ocl_ximage_sum": ([P, P, P, I, I, LG], I), "ocl_complex_abs": ([P, P, P, LG, I], I),
        "ocl_copy": ([P, P, P, LG], I), "ocl_fft2d": ([P, P, P, LG, I, I, I], I),
        "ocl_simple_mri_recon": ([P, P, P, P, I, I, I, I], I),
        "ocl_simple_mri_recon_batch": ([P, I, P, P, P, I, I, I, I], I),
        "ocl_scal": ([P, P, LG, F32, F32], I), "ocl_axpy": ([P, LG, F32, F32, P, P], I),
        "ocl_reduce": ([P, P, P, LG, I, ctypes.POINTER(D)], I),
        "ocl_normest": ([P, I, LG, ctypes.POINTER(D), ctypes.POINTER(I)], I),
        "ocl_ctx_set_normest_stream": ([P, I], I),
        "ocl_ctx_normest_stream_position": ([P, ctypes.POINTER(ctypes.c_longlong)], I),
        "ocl_nesta_run": ([P, P, P, P, P, I, I, I, I, ctypes.POINTER(NestaParams), I, ctypes.POINTER(NestaReport)],
                          I),
        "ocl_nesta_run_noalloc": ([P, P, P, P, P, I, I, I, I, ctypes.POINTER(NestaParams), I,
                                   ctypes.POINTER(NestaReport)], I),
        "ocl_motion_compensation": ([P, P, P, I, I, I, ctypes.POINTER(McArgs)], I),
        "ocl_adjoint_motion_compensation": ([P, P, P, I, I, I, ctypes.POINTER(McArgs)], I),
        "ocl_nesta_run_mc": ([P, P, P, P, P, I, I, I, I, ctypes.POINTER(NestaParams), I, ctypes.POINTER(McArgs),
                              ctypes.POINTER(NestaReport)], I),
        "ocl_gwr_create": ([P, I, I, I, ctypes.POINTER(GwrParams), PP], I),
        "ocl_gwr_info": ([P, ctypes.POINTER(GwrInfo)], I),
        "ocl_gwr_run": ([P, P, P, I, P, P, ctypes.POINTER(GwrReport)], I),
        "ocl_gwr_destroy": ([P], I),
        "ocl_batch_create": ([P, I, ctypes.POINTER(Slice), ctypes.POINTER(NestaParams), PP], I),
        "ocl_batch_create_mc": ([P, I, ctypes.POINTER(Slice), ctypes.POINTER(NestaParams), ctypes.POINTER(McArgs), PP],
                                I),
        "ocl_batch_create_form": ([P, I, ctypes.POINTER(Slice), ctypes.POINTER(NestaParams), I, PP], I),
        "ocl_ctx_set_autotune": ([P, I], I),
        "ocl_ctx_set_split_gram": ([P, I], I), "ocl_ctx_set_group_threshold": ([P, I], I),
        "ocl_rsos": ([P, P, P, I, I, LG], I), "ocl_simple_mri_recon_sos": ([P, P, P, I, I, I, I], I),
        "ocl_ctx_set_tiles": ([P, I, I, I, I, I, I], I),
        "ocl_autotune_shape": ([P, I, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I),
                                ctypes.POINTER(D), ctypes.POINTER(D)], I),
        "ocl_batch_setup": ([P], I), "ocl_batch_iterate": ([P, I], I), "ocl_batch_capture": ([P, I], I),
        "ocl_batch_replay": ([P], I), "ocl_batch_active": ([P, ctypes.POINTER(I)], I),
        "ocl_batch_finalize": ([P], I), "ocl_batch_report": ([P, I, ctypes.POINTER(NestaReport)], I),
        "ocl_batch_run": ([P, I], I), "ocl_batch_destroy": ([P], I),
        "ocl_batch_info": ([P, ctypes.POINTER(I), ctypes.POINTER(LG), ctypes.POINTER(LG)], I),
        "ocl_batch_time": ([P, I, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)], I),
        "ocl_timer_begin": ([P, PP], I), "ocl_timer_end": ([P, P, ctypes.POINTER(D)], I),
        "ocl_add_noise_masked": ([P, P, P, I, I, LG, F32, ctypes.c_ulonglong], I),
        "ocl_server_create": ([I, ctypes.c_size_t, I, PP], I),
        "ocl_server_submit": ([P, ctypes.POINTER(Slice), ctypes.POINTER(NestaParams), ctypes.POINTER(NestaReport),
                               ctypes.POINTER(LG)], I),
        "ocl_server_poll": ([P, LG, ctypes.POINTER(I), ctypes.POINTER(D)], I),
        "ocl_server_wait": ([P, LG, D, ctypes.POINTER(I)], I),
        "ocl_server_request_times": ([P, LG, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)], I),
        "ocl_server_get_stats": ([P, ctypes.POINTER(ServerStats)], I),
        "ocl_server_inject_failure": ([P, ctypes.c_char_p], I),
        "ocl_server_destroy": ([P], I),
    }
    for name, (a, r) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise OclError(rc, lib().ocl_last_error().decode(errors="replace"))


class DeviceArray:
    """A device allocation with a numpy-like shape/dtype tag."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().ocl_malloc(ctx.handle, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().ocl_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, "size mismatch"
        _check(lib().ocl_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def free(self):
        if self.ptr and self.ctx.handle:
            lib().ocl_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device + one stream (CLapp)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().ocl_ctx_create(device, ctypes.byref(h)))
        self.handle = h.value
        self._children = weakref.WeakSet()      # batches and registrations: destroyed before the context

    def close(self):
        if self.handle:
            for ch in list(self._children):
                ch.close()
            lib().ocl_ctx_destroy(self.handle)
            self.handle = None

    def sync(self):
        _check(lib().ocl_ctx_sync(self.handle))

    def info(self):
        name = ctypes.create_string_buffer(256)
        mem = ctypes.c_size_t()
        cus = ctypes.c_int()
        _check(lib().ocl_device_info(self.handle, name, 256, ctypes.byref(mem), ctypes.byref(cus)))
        return dict(name=name.value.decode(), total_mem=mem.value, num_cus=cus.value)

    def timer_begin(self):
        """Records a start event on the context's stream; pass the handle to timer_end."""
        h = ctypes.c_void_p()
        _check(lib().ocl_timer_begin(self.handle, ctypes.byref(h)))
        return h

    def timer_end(self, h):
        """Records the stop event, waits for it and returns the elapsed device time in ms."""
        ms = ctypes.c_double()
        _check(lib().ocl_timer_end(self.handle, h, ctypes.byref(ms)))
        return ms.value

    def set_autotune(self, on=True):
        """With autotuning on, batches created on this context tune the tile shape of each new H x W at setup."""
        _check(lib().ocl_ctx_set_autotune(self.handle, 1 if on else 0))

    def set_split_gram(self, on=True):
        """Split-precision form: Gram pass for line masks where it is instantiated (True) or row and column passes."""
        _check(lib().ocl_ctx_set_split_gram(self.handle, 1 if on else 0))

    def set_group_threshold(self, min_workgroups):
        """Column-pass workgroups from which a size of a mixed table gets launches of its own (tables built later)."""
        _check(lib().ocl_ctx_set_group_threshold(self.handle, int(min_workgroups)))

    def set_tiles(self, H, W, ca_fast, tile_cols, b_fast, rows_per_wg):
        """Pins the tile shape of the FP32 passes for H x W."""
        _check(lib().ocl_ctx_set_tiles(self.handle, int(H), int(W), int(ca_fast), int(tile_cols), int(b_fast),
                                       int(rows_per_wg)))

    def autotune_shape(self, H, W):
        """Tunes (or returns the cached) tile shape of the FP32 passes for H x W."""
        v = [ctypes.c_int() for _ in range(4)]
        t = [ctypes.c_double() for _ in range(2)]
        _check(lib().ocl_autotune_shape(self.handle, int(H), int(W), *[ctypes.byref(x) for x in v],
                                        *[ctypes.byref(x) for x in t]))
        return dict(ca_fast=v[0].value, tile_cols=v[1].value, b_fast=v[2].value, rows_per_wg=v[3].value,
                    ms_CA=t[0].value, ms_B=t[1].value)

    # --- memory ---
    def empty(self, shape, dtype=np.complex64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def pinned(self, a: np.ndarray) -> np.ndarray:
        """A copy of `a` in page-locked host memory (ocl_host_malloc), as a numpy array.  Uploads from it are one
        DMA transfer.  The memory is released when the array and every view of it are gone."""
        a = np.ascontiguousarray(a)
        n = max(a.nbytes, 1)
        p = ctypes.c_void_p()
        _check(lib().ocl_host_malloc(self.handle, n, ctypes.byref(p)))
        buf = type("PinnedBuffer", (ctypes.c_char * n,), {}).from_address(p.value)
        # the numpy array keeps `buf` alive; the block is freed (no context needed) once `buf` is collected, and is
        # left to the process's end rather than freed while the runtime shuts down
        fin = weakref.finalize(buf, lib().ocl_host_free, None, p.value)
        fin.atexit = False
        h = np.frombuffer(buf, dtype=a.dtype, count=a.size).reshape(a.shape)
        h[...] = a
        return h

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        return DeviceArray(self, a.shape, a.dtype).set(a)

    def zeros(self, shape, dtype=np.complex64) -> DeviceArray:
        d = DeviceArray(self, shape, dtype)
        _check(lib().ocl_memset_zero(self.handle, d.ptr, d.nbytes))
        return d

    # --- Processes ---
    def ttv(self, x: DeviceArray, out: DeviceArray, forward=True):
        F = x.shape[0]
        _check(lib().ocl_ttv(self.handle, x.ptr, out.ptr, F, x.size // F, 1 if forward else 0))

    def vecnorm(self, x, out, mu):
        _check(lib().ocl_vecnorm(self.handle, x.ptr, out.ptr, x.size, float(mu)))

    def cep(self, x, S, out, F, C, in_coils, conj):
        _check(lib().ocl_cep(self.handle, x.ptr, S.ptr, out.ptr, F, C, S.size // C, int(in_coils), int(conj)))

    def apply_mask(self, k, M):
        F, C = k.shape[:2]
        _check(lib().ocl_apply_mask(self.handle, k.ptr, M.ptr, F, C, k.size // (F * C)))

    def ximage_sum(self, k, out):
        F, C = k.shape[:2]
        _check(lib().ocl_ximage_sum(self.handle, k.ptr, out.ptr, F, C, k.size // (F * C)))

    def complex_abs(self, x, out, pow2=False):
        _check(lib().ocl_complex_abs(self.handle, x.ptr, out.ptr, x.size, int(pow2)))

    def copy(self, x, out):
        _check(lib().ocl_copy(self.handle, x.ptr, out.ptr, x.size))

    def fft2d(self, x, out, inverse=False):
        H, W = x.shape[-2:]
        _check(lib().ocl_fft2d(self.handle, x.ptr, out.ptr, x.size // (H * W), H, W, 1 if inverse else -1))

    def simple_mri_recon(self, y, S, x):
        F, C, H, W = y.shape
        _check(lib().ocl_simple_mri_recon(self.handle, y.ptr, S.ptr, x.ptr, F, C, H, W))

    def simple_mri_recon_batch(self, items):
        """SimpleMRIRecon of many inputs of one shape at once: items = [(y, S, x), ...] DeviceArrays."""
        n = len(items)
        if n == 0:
            return
        F, C, H, W = items[0][0].shape
        arr = lambda k: (ctypes.c_void_p * n)(*[it[k].ptr for it in items])
        _check(lib().ocl_simple_mri_recon_batch(self.handle, n, arr(0), arr(1), arr(2), F, C, H, W))

    def rsos(self, k, out):
        F, C = k.shape[:2]
        _check(lib().ocl_rsos(self.handle, k.ptr, out.ptr, F, C, k.size // (F * C)))

    def simple_mri_recon_sos(self, y, x):
        F, C, H, W = y.shape
        _check(lib().ocl_simple_mri_recon_sos(self.handle, y.ptr, x.ptr, F, C, H, W))

    def scal(self, x, a):
        a = complex(a)
        _check(lib().ocl_scal(self.handle, x.ptr, x.size, a.real, a.imag))

    def axpy(self, a, x, y):
        a = complex(a)
        _check(lib().ocl_axpy(self.handle, x.size, a.real, a.imag, x.ptr, y.ptr))

    def reduce(self, x, op, y=None):
        r = (ctypes.c_double * 2)()
        _check(lib().ocl_reduce(self.handle, x.ptr, y.ptr if y is not None else None, x.size, RED[op], r))
        return r[0], r[1]

    def add_noise_masked(self, y, M, sigma, seed):
        F, C = y.shape[:2]
        _check(lib().ocl_add_noise_masked(self.handle, y.ptr, M.ptr, F, C, y.size // (F * C), float(sigma),
                                          int(seed) & ((1 << 64) - 1)))

    def set_normest_stream(self, carry=True):
        """carry: one drand48 stream across this context's normests (the reference's process-wide stream), reset to
        the default state; False: a fresh stream per estimate."""
        _check(lib().ocl_ctx_set_normest_stream(self.handle, 1 if carry else 0))

    def normest_stream_position(self):
        v = ctypes.c_longlong()
        _check(lib().ocl_ctx_normest_stream_position(self.handle, ctypes.byref(v)))
        return v.value

    def normest(self, F, N):
        v = ctypes.c_double()
        it = ctypes.c_int()
        _check(lib().ocl_normest(self.handle, F, N, ctypes.byref(v), ctypes.byref(it)))
        return v.value, it.value

    # --- NESTA ---
    def nesta(self, y: DeviceArray, S: DeviceArray, M: DeviceArray, x: DeviceArray, prm: NestaParams | None = None,
              form=OCL_FORM_FUSED):
        prm = prm or nesta_params()
        F, C, H, W = y.shape
        rep = NestaReport()
        _check(lib().ocl_nesta_run(self.handle, y.ptr, S.ptr, M.ptr, x.ptr, F, C, H, W, ctypes.byref(prm), form,
                                   ctypes.byref(rep)))
        return rep.as_dict()

    def nesta_noalloc(self, y: DeviceArray, S: DeviceArray, M: DeviceArray, x: DeviceArray,
                      prm: NestaParams | None = None, form=OCL_FORM_FUSED):
        """ocl_nesta_run_noalloc: the same reconstruction with every work buffer in the device's static scratch."""
        prm = prm or nesta_params()
        F, C, H, W = y.shape
        rep = NestaReport()
        _check(lib().ocl_nesta_run_noalloc(self.handle, y.ptr, S.ptr, M.ptr, x.ptr, F, C, H, W, ctypes.byref(prm),
                                           form, ctypes.byref(rep)))
        return rep.as_dict()

    # --- motion compensation ---
    @staticmethod
    def mc_args(xn: DeviceArray, mesh: DeviceArray, bound_box) -> McArgs:
        a = McArgs(xn.ptr, mesh.ptr)
        for i, v in enumerate(bound_box):
            a.bound_box[i] = int(v)
        return a

    def motion_compensation(self, x: DeviceArray, out: DeviceArray, mc: McArgs, adjoint=False):
        F, H, W = x.shape
        fn = lib().ocl_adjoint_motion_compensation if adjoint else lib().ocl_motion_compensation
        _check(fn(self.handle, x.ptr, out.ptr, F, H, W, ctypes.byref(mc)))

    def nesta_mc(self, y: DeviceArray, S: DeviceArray, M: DeviceArray, x: DeviceArray, mc: McArgs,
                 prm: NestaParams | None = None, form=OCL_FORM_PLAIN):
        """NESTA with motion compensation: x holds x_ref on entry (the previous reconstruction)."""
        prm = prm or nesta_params()
        F, C, H, W = y.shape
        rep = NestaReport()
        _check(lib().ocl_nesta_run_mc(self.handle, y.ptr, S.ptr, M.ptr, x.ptr, F, C, H, W, ctypes.byref(prm), form,
                                      ctypes.byref(mc), ctypes.byref(rep)))
        return rep.as_dict()

    def registration(self, F, H, W, prm: GwrParams | None = None) -> "Registration":
        return Registration(self, F, H, W, prm or gwr_params())

    def batch(self, slices, prm: NestaParams | None = None, form=OCL_FORM_FUSED, mcs=None) -> "Batch":
        return Batch(self, slices, prm or nesta_params(), form, mcs)


class Batch:
    """Many uneven slices advanced together (ocl_batch_*).  `slices` = list of (y, S, M, x_out) DeviceArrays."""

    def __init__(self, ctx: Context, slices, prm: NestaParams, form=OCL_FORM_FUSED, mcs=None):
        """mcs: one McArgs per slice for motion compensation (fused form; each x_out then holds x_ref)."""
        self.ctx = ctx
        self.handle = None
        self.keep = list(slices)
        arr = (Slice * len(slices))()
        for i, (y, S, M, x) in enumerate(slices):
            F, C, H, W = y.shape
            arr[i] = Slice(y.ptr, S.ptr, M.ptr, x.ptr, F, C, H, W)
        self.prm = prm
        h = ctypes.c_void_p()
        if mcs is None:
            _check(lib().ocl_batch_create_form(ctx.handle, len(slices), arr, ctypes.byref(prm), int(form),
                                               ctypes.byref(h)))
        else:
            marr = (McArgs * len(mcs))(*mcs)
            _check(lib().ocl_batch_create_mc(ctx.handle, len(slices), arr, ctypes.byref(prm), marr, ctypes.byref(h)))
        self.handle = h.value
        ctx._children.add(self)

    def setup(self):
        _check(lib().ocl_batch_setup(self.handle))

    def iterate(self, n):
        _check(lib().ocl_batch_iterate(self.handle, int(n)))

    def capture(self, n):
        _check(lib().ocl_batch_capture(self.handle, int(n)))

    def replay(self):
        _check(lib().ocl_batch_replay(self.handle))

    def active(self):
        n = ctypes.c_int()
        _check(lib().ocl_batch_active(self.handle, ctypes.byref(n)))
        return n.value

    def finalize(self):
        _check(lib().ocl_batch_finalize(self.handle))

    def run(self, chunk=16):
        _check(lib().ocl_batch_run(self.handle, int(chunk)))

    def time_iters(self, n):
        """Runs n iterations with HIP events around each launch: (ms of pass B, ms of pass CA, total ms)."""
        b, c, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().ocl_batch_time(self.handle, int(n), ctypes.byref(b), ctypes.byref(c), ctypes.byref(t)))
        return b.value, c.value, t.value

    def report(self, i):
        rep = NestaReport()
        _check(lib().ocl_batch_report(self.handle, i, ctypes.byref(rep)))
        return rep.as_dict()

    def info(self):
        g, b, c = ctypes.c_int(), ctypes.c_long(), ctypes.c_long()
        _check(lib().ocl_batch_info(self.handle, ctypes.byref(g), ctypes.byref(b), ctypes.byref(c)))
        return {"launch_groups": g.value, "wg_B": b.value, "wg_CA": c.value}

    def close(self):
        if self.handle:
            lib().ocl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching NESTA-TV server (ocl_server_*).  It runs its own stream on `device`; inputs/outputs are
    DeviceArrays of any Context on the same device and must outlive their request."""

    def __init__(self, device: int = 0, budget_bytes: int = 64 << 30, chunk: int = 8):
        h = ctypes.c_void_p()
        _check(lib().ocl_server_create(device, int(budget_bytes), int(chunk), ctypes.byref(h)))
        self.handle = h.value
        self.keep = {}

    def submit(self, y, S, M, x, prm: NestaParams, want_report=False):
        F, C, H, W = y.shape
        sl = Slice(y.ptr, S.ptr, M.ptr, x.ptr, F, C, H, W)
        rep = NestaReport() if want_report else None
        rid = ctypes.c_long()
        _check(lib().ocl_server_submit(self.handle, ctypes.byref(sl), ctypes.byref(prm),
                                       ctypes.byref(rep) if rep is not None else None, ctypes.byref(rid)))
        self.keep[rid.value] = (y, S, M, x, prm, rep)
        return rid.value

    def poll(self, rid):
        st, lat = ctypes.c_int(), ctypes.c_double()
        _check(lib().ocl_server_poll(self.handle, rid, ctypes.byref(st), ctypes.byref(lat)))
        return st.value, lat.value

    def wait(self, rid, timeout=600.0):
        st = ctypes.c_int()
        _check(lib().ocl_server_wait(self.handle, rid, float(timeout), ctypes.byref(st)))
        return st.value

    def times(self, rid):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().ocl_server_request_times(self.handle, rid, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def report(self, rid):
        rep = self.keep[rid][5]
        return rep.as_dict() if rep is not None else None

    def inject_failure(self, message: str):
        _check(lib().ocl_server_inject_failure(self.handle, message.encode()))

    def stats(self):
        st = ServerStats()
        _check(lib().ocl_server_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.handle:
            lib().ocl_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Registration:
    """GroupwiseRegistration on the device (ocl_gwr_*): init() at construction, run() = launch()."""

    def __init__(self, ctx: Context, F, H, W, prm: GwrParams):
        self.ctx = ctx
        self.handle = None
        self.F, self.H, self.W = F, H, W
        h = ctypes.c_void_p()
        _check(lib().ocl_gwr_create(ctx.handle, F, H, W, ctypes.byref(prm), ctypes.byref(h)))
        self.handle = h.value
        ctx._children.add(self)
        info = GwrInfo()
        _check(lib().ocl_gwr_info(self.handle, ctypes.byref(info)))
        self.Cp, self.support, self.box, self.areaX = info.Cp, info.support, info.box, info.areaX
        self.use_lambda = bool(info.use_lambda)
        self.bound_box = [int(v) for v in info.bound_box]
        self.mesh_ptr = info.mesh

    def mesh(self) -> np.ndarray:
        """The original mesh x, int32 [2][W][H] (1-based)."""
        out = np.empty((2, self.W, self.H), np.int32)
        _check(lib().ocl_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.mesh_ptr, out.nbytes))
        return out

    def mc_args(self, xn: DeviceArray) -> McArgs:
        a = McArgs(xn.ptr, self.mesh_ptr)
        for i, v in enumerate(self.bound_box):
            a.bound_box[i] = v
        return a

    def run(self, x: DeviceArray, T: DeviceArray, xn: DeviceArray, T_in=False, registered: DeviceArray | None = None):
        rep = GwrReport()
        _check(lib().ocl_gwr_run(self.handle, x.ptr, T.ptr, int(bool(T_in)), xn.ptr,
                                 registered.ptr if registered is not None else None, ctypes.byref(rep)))
        return rep.as_dict()

    def close(self):
        if self.handle:
            lib().ocl_gwr_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
