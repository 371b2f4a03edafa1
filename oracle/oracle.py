"""ORACLE — test infrastructure only.

ctypes front end of oracle/liboracle.so, the CPU restatement of OpenCLIPER's NESTA-TV path (see the header of
oracle/nesta_oracle.cpp for the reference file:line each function follows and for how each one is pinned).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MAX_STAGES = 16
TRACE = 4096


class Params(ctypes.Structure):
    """NestaUp::LaunchParameters field order (include/OpenCLIPER/processes/nesta/NestaUp.hpp:67-83)."""
    _fields_ = [("lambda_i", ctypes.c_double), ("mu_f", ctypes.c_double), ("La", ctypes.c_double),
                ("maxIntIter", ctypes.c_int), ("tolVar", ctypes.c_double), ("verbose", ctypes.c_int),
                ("maxIter", ctypes.c_int), ("stoptest", ctypes.c_int), ("miniter", ctypes.c_int)]


class Report(ctypes.Structure):
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
        d["trace_f"] = np.array(self.trace_f[: self.ntrace])
        d["trace_l2"] = np.array(self.trace_l2[: self.ntrace])
        d["trace_l1"] = np.array(self.trace_l1[: self.ntrace])
        return d


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P, I, LG, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        sig = {
            "orc_fft2d_f32": ([P, P, LG, I, I, I], None), "orc_fft2d_f64": ([P, P, LG, I, I, I], None),
            "orc_ttv_f32": ([P, P, I, LG, I], None), "orc_ttv_f64": ([P, P, I, LG, I], None),
            "orc_vecnorm_f32": ([P, P, LG, ctypes.c_float], None),
            "orc_cep_f32": ([P, P, P, I, I, LG, I, I], None),
            "orc_apply_mask_f32": ([P, P, I, I, LG], None),
            "orc_ximage_sum_f32": ([P, P, I, I, LG], None),
            "orc_nrm2_f32": ([P, LG], D), "orc_dotc_re_f32": ([P, P, LG], D), "orc_asum_f32": ([P, LG], D),
            "orc_op_A_f64": ([P, P, P, P, I, I, I, I], None), "orc_op_At_f64": ([P, P, P, I, I, I, I], None),
            "orc_normest_f32": ([I, LG, ctypes.POINTER(I)], D), "orc_normest_f64": ([I, LG, ctypes.POINTER(I)], D),
            "orc_drand48_nth": ([LG], D), "orc_drand48_nth_seek": ([LG], D),
            "orc_drand48_seek": ([ctypes.c_longlong], None), "orc_drand48_position": ([], ctypes.c_longlong),
            "orc_nesta_f32": ([P, P, P, P, I, I, I, I, ctypes.POINTER(Params), ctypes.POINTER(Report)], I),
            "orc_nesta_f64": ([P, P, P, P, I, I, I, I, ctypes.POINTER(Params), ctypes.POINTER(Report)], I),
            "orc_simple_recon_f32": ([P, P, P, I, I, I, I], I), "orc_simple_recon_f64": ([P, P, P, I, I, I, I], I),
            "orc_rsos_f32": ([P, P, I, I, LG], None), "orc_simple_recon_sos_f64": ([P, P, I, I, I, I], I),
            "orc_num_threads": ([], I), "orc_set_threads": ([I], None),
            "orc_stop_rule_f64": ([P, I, I, D, I, I, P], I),
            "orc_mc_f64": ([P, P, P, P, P, I, I, I, I], I), "orc_mc_f32": ([P, P, P, P, P, I, I, I, I], I),
            "orc_normest_mc_f64": ([P, P, P, I, I, I, ctypes.POINTER(I)], D),
            "orc_nesta_mc_f64": ([P, P, P, P, I, I, I, I, P, P, P, ctypes.POINTER(Params), ctypes.POINTER(Report)], I),
            "orc_nesta_mc_f32": ([P, P, P, P, I, I, I, I, P, P, P, ctypes.POINTER(Params), ctypes.POINTER(Report)], I),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def params(lambda_i=1e-2, mu_f=3e-4, La=1.0, maxIntIter=4, tolVar=3e-4, verbose=1, maxIter=100, stoptest=1,
           miniter=7) -> Params:
    """Defaults are the ones tests/MRIRecon.cpp:105-114 passes to NestaUp."""
    return Params(lambda_i, mu_f, La, maxIntIter, tolVar, verbose, maxIter, stoptest, miniter)


def fft2d(x, inverse=False, dtype=np.complex128):
    a = _c(x, dtype)
    out = np.empty_like(a)
    H, W = a.shape[-2:]
    batch = a.size // (H * W)
    f = lib().orc_fft2d_f64 if dtype == np.complex128 else lib().orc_fft2d_f32
    f(_p(a), _p(out), batch, H, W, 1 if inverse else -1)
    return out


def ttv(x, forward=True, dtype=np.complex128):
    a = _c(x, dtype)
    out = np.empty_like(a)
    F = a.shape[0]
    f = lib().orc_ttv_f64 if dtype == np.complex128 else lib().orc_ttv_f32
    f(_p(a), _p(out), F, a.size // F, 1 if forward else 0)
    return out


def vecnorm(x, mu):
    a = _c(x, np.complex64)
    out = np.empty_like(a)
    lib().orc_vecnorm_f32(_p(a), _p(out), a.size, mu)
    return out


def cep(x, S, F, C, in_coils, conj):
    a, s = _c(x, np.complex64), _c(S, np.complex64)
    N = s.size // C
    out = np.empty((F, C) + s.shape[1:], dtype=np.complex64)
    lib().orc_cep_f32(_p(a), _p(s), _p(out), F, C, N, int(in_coils), int(conj))
    return out


def apply_mask(k, M):
    a = _c(k, np.complex64).copy()
    m = _c(M, np.uint8)
    F, C = a.shape[:2]
    lib().orc_apply_mask_f32(_p(a), _p(m), F, C, a.size // (F * C))
    return a


def ximage_sum(k):
    a = _c(k, np.complex64)
    F, C = a.shape[:2]
    out = np.empty((F,) + a.shape[2:], dtype=np.complex64)
    lib().orc_ximage_sum_f32(_p(a), _p(out), F, C, a.size // (F * C))
    return out


def op_A(x, S, M):
    x, S, M = _c(x, np.complex128), _c(S, np.complex128), _c(M, np.uint8)
    F, H, W = x.shape
    C = S.shape[0]
    out = np.empty((F, C, H, W), dtype=np.complex128)
    lib().orc_op_A_f64(_p(x), _p(S), _p(M), _p(out), F, C, H, W)
    return out


def op_At(r, S):
    r, S = _c(r, np.complex128), _c(S, np.complex128)
    F, C, H, W = r.shape
    out = np.empty((F, H, W), dtype=np.complex128)
    lib().orc_op_At_f64(_p(r), _p(S), _p(out), F, C, H, W)
    return out


def drand48_position():
    """Draws taken from the restatement's drand48 stream (the reference's unseeded, process-wide drand48)."""
    return int(lib().orc_drand48_position())


def normest(F, N, double=True, drand_start=0):
    """myNormest; drand_start: draws already taken from the process's drand48 stream when it runs (0: the first
    normest of a process), None: continue where the last call left it."""
    if drand_start is not None:
        lib().orc_drand48_seek(int(drand_start))
    it = ctypes.c_int(0)
    f = lib().orc_normest_f64 if double else lib().orc_normest_f32
    v = f(F, N, ctypes.byref(it))
    return v, it.value


def nesta(y, S, M, prm: Params | None = None, double=True, drand_start=0):
    """Runs the NESTA-TV reconstruction; returns (x, report dict).  drand_start as for normest; the report's
    "drand48_position" is where the stream is after the call."""
    if drand_start is not None:
        lib().orc_drand48_seek(int(drand_start))
    prm = prm or params()
    dt = np.complex128 if double else np.complex64
    y, S, M = _c(y, dt), _c(S, dt), _c(M, np.uint8)
    F, C, H, W = y.shape
    x = np.empty((F, H, W), dtype=dt)
    rep = Report()
    f = lib().orc_nesta_f64 if double else lib().orc_nesta_f32
    rc = f(_p(y), _p(S), _p(M), _p(x), F, C, H, W, ctypes.byref(prm), ctypes.byref(rep))
    if rc != 0:
        raise RuntimeError(f"oracle nesta failed ({rc})")
    out = rep.as_dict()
    out["drand48_position"] = drand48_position()
    return x, out


def simple_recon(y, S, double=True):
    dt = np.complex128 if double else np.complex64
    y, S = _c(y, dt), _c(S, dt)
    F, C, H, W = y.shape
    x = np.empty((F, H, W), dtype=dt)
    f = lib().orc_simple_recon_f64 if double else lib().orc_simple_recon_f32
    f(_p(y), _p(S), _p(x), F, C, H, W)
    return x


def rsos(k):
    """rss.cl: (sqrt(sum_c |k[f][c]|^2), 0) per pixel, FP32 with the coil sum in order."""
    a = _c(k, np.complex64)
    F, C = a.shape[:2]
    out = np.empty((F,) + a.shape[2:], dtype=np.complex64)
    lib().orc_rsos_f32(_p(a), _p(out), F, C, a.size // (F * C))
    return out


def simple_recon_sos(y):
    """SimpleMRIReconSOS in FP64: inverse 2-D FFT of each coil image, then the root sum of squares."""
    y = np.array(y, dtype=np.complex128)
    F, C, H, W = y.shape
    x = np.empty((F, H, W), dtype=np.complex128)
    lib().orc_simple_recon_sos_f64(_p(y), _p(x), F, C, H, W)
    return x


def stop_rule(f, miniter, tolVar, maxIter, nstages=1):
    """The stop test (stoptest = 1) on a scripted objective sequence; returns the iterations of each stage."""
    a = _c(f, np.float64)
    iters = np.zeros(nstages, np.int32)
    used = lib().orc_stop_rule_f64(_p(a), a.size, miniter, tolVar, maxIter, nstages, _p(iters))
    if used < 0:
        raise ValueError("objective sequence ran out")
    return [int(v) for v in iters]


def mc(x, xn, mesh, bb, adjoint=False, double=True):
    """MotionCompensation (adjoint=False) or AdjointMotionCompensation of x [F][H][W]; xn [2][F][H][W] float32
    sample positions (1-based), mesh [2][W][H] int32 original mesh (1-based), bb the 1-based bound box."""
    dt = np.complex128 if double else np.complex64
    a = _c(x, dt)
    F, H, W = a.shape
    xn, mesh, bb = _c(xn, np.float32), _c(mesh, np.int32), _c(bb, np.uint32)
    assert xn.shape == (2, F, H, W) and mesh.shape == (2, W, H) and bb.shape == (4,)
    out = np.empty_like(a)
    f = lib().orc_mc_f64 if double else lib().orc_mc_f32
    if f(_p(a), _p(out), _p(xn), _p(mesh), _p(bb), F, H, W, int(adjoint)) != 0:
        raise ValueError("motion compensation box leaves the frame")
    return out


def normest_mc(xn, mesh, bb, drand_start=0):
    """myNormest of U = tTV o MC in FP64 (drand_start as for normest)."""
    if drand_start is not None:
        lib().orc_drand48_seek(int(drand_start))
    xn, mesh, bb = _c(xn, np.float32), _c(mesh, np.int32), _c(bb, np.uint32)
    _, F, H, W = xn.shape
    it = ctypes.c_int(0)
    v = lib().orc_normest_mc_f64(_p(xn), _p(mesh), _p(bb), F, H, W, ctypes.byref(it))
    if v < 0:
        raise ValueError("motion compensation box leaves the frame")
    return v, it.value


def nesta_mc(y, S, M, x_ref, xn, mesh, bb, prm: Params | None = None, double=True, drand_start=0):
    """NESTA with motion compensation (argsMC != nullptr), starting from x_ref; returns (x, report dict).  drand_start
    as for nesta."""
    if drand_start is not None:
        lib().orc_drand48_seek(int(drand_start))
    prm = prm or params()
    dt = np.complex128 if double else np.complex64
    y, S, M = _c(y, dt), _c(S, dt), _c(M, np.uint8)
    F, C, H, W = y.shape
    x = np.array(x_ref, dtype=dt)
    xn, mesh, bb = _c(xn, np.float32), _c(mesh, np.int32), _c(bb, np.uint32)
    rep = Report()
    f = lib().orc_nesta_mc_f64 if double else lib().orc_nesta_mc_f32
    rc = f(_p(y), _p(S), _p(M), _p(x), F, C, H, W, _p(xn), _p(mesh), _p(bb), ctypes.byref(prm), ctypes.byref(rep))
    if rc != 0:
        raise RuntimeError(f"oracle nesta_mc failed ({rc})")
    out = rep.as_dict()
    out["drand48_position"] = drand48_position()
    return x, out


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
