"""Writes the NESTA-TV golden fixtures under tests/golden/ from an independent numpy FP64 implementation.

This is a second, separately written statement of OpenCLIPER's NestaUp::launch
(src/processes/nesta/NestaUp.cpp:442-910) and myNormest (:291-439): vectorised numpy with numpy.fft, its own
drand48 generator, its own stop-test bookkeeping.  It shares no code with oracle/nesta_oracle.cpp, so agreement
between the two pins the oracle's NESTA (tests/test_oracle.py::test_nesta_golden).

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from opencliper_amd import synth  # noqa: E402  (input generator only)

LONG_MAX = float(2 ** 63 - 1)


class Drand48:
    """POSIX drand48 from its default state (X0 = 0x1234ABCD330E, a = 0x5DEECE66D, c = 0xB, mod 2^48)."""

    def __init__(self):
        self.x = 0x1234ABCD330E

    def take(self, n):
        out = np.empty(n)
        x = self.x
        for i in range(n):
            x = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
            out[i] = x / float(1 << 48)
        self.x = x
        return out


def D(x):
    """Periodic backward difference over frames (axis 0)."""
    return x - np.roll(x, 1, axis=0)


def Dt_neg(v):
    """The reference's 'adjoint' temporal TV kernel, which is -D^T: out_f = v_{f+1} - v_f."""
    return np.roll(v, -1, axis=0) - v


def power_norm(F, HW):
    x = np.ones((F, HW), dtype=np.complex128)
    e = np.linalg.norm(x)
    x = x / e
    e0 = 0.0
    cnt = 0
    rng = Drand48()
    while abs(e - e0) > 1e-3 * e and cnt < 100:
        e0 = e
        Sx = D(x)
        if np.sum(np.abs(Sx.real) + np.abs(Sx.imag)) == 0:
            Sx = rng.take(F * HW).reshape(F, HW).astype(np.complex128)
        e = np.linalg.norm(Sx)
        x = Dt_neg(Sx)
        x = x / np.linalg.norm(x)
        cnt += 1
    return e, cnt


def nesta_numpy(y, S, M, lambda_i=1e-2, mu_f=3e-4, La=1.0, maxIntIter=4, tolVar=3e-4, maxIter=100, stoptest=1,
                miniter=7):
    y = y.astype(np.complex128)
    S = S.astype(np.complex128)
    F, C, H, W = y.shape
    N = H * W
    m = (M != 0)[:, None, :, :]

    def A_unit(x):      # mask * FFT2(S x) / sqrt(N)
        return m * np.fft.fft2(S[None] * x[:, None], axes=(-2, -1)) / np.sqrt(N)

    def AH_unit(r):     # sum_c conj(S) IFFT2(r) * sqrt(N)   (unitary adjoint, mask not re-applied)
        return np.sum(np.conj(S)[None] * np.fft.ifft2(r, axes=(-2, -1)), axis=1) * np.sqrt(N)

    xref = AH_unit(y)
    lam = lambda_i * np.abs(xref).max()
    muf = mu_f * np.abs(xref).max()
    mu0 = 0.9 * np.abs(D(xref)).max()
    normU, ne_it = power_norm(F, N)
    if lam / La > mu0:
        mu0 = lam / La
    gamma = (muf / mu0) ** (1.0 / maxIntIter)
    gammat = (tolVar / 0.1) ** (1.0 / maxIntIter)
    tv = 0.1
    mu = mu0
    iters, trace = [], []
    fmean = np.zeros(miniter)
    for _stage in range(maxIntIter):
        mu *= gamma
        tv *= gammat
        Lmu = lam * normU / mu + La
        Wk = np.zeros_like(xref)
        xk = xref.copy()
        fmean[miniter - 1] = LONG_MAX
        qp_hit = False
        k = 0
        n_done = 0
        while k < maxIter:
            n_done += 1
            Ux = D(xk)
            u = Ux / np.maximum(mu, np.abs(Ux))
            fx = np.real(np.vdot(u, Ux))
            r = A_unit(xk) - y
            l2 = np.sum(np.abs(r) ** 2)
            df = lam * (-Dt_neg(u)) + AH_unit(r)
            yk = xk - df / Lmu
            f = 0.5 * l2 + lam * (fx - mu / 2 * np.linalg.norm(u) ** 2)
            trace.append(f)
            if stoptest == 1:
                write = True
                if k >= miniter:
                    mean = np.sum(fmean) / miniter
                    qp = (mean - f) / mean
                    if qp <= tv:
                        if qp_hit:
                            k = maxIter
                        qp_hit = True
                        write = False
                if write:
                    fmean[(miniter - 1) - ((k + 1) % miniter)] = f
            apk = 0.5 * (k + 1)
            tau = 2.0 / (k + 3)
            Wk = Wk + apk * df
            zk = xref - Wk / Lmu
            xk = tau * zk + (1 - tau) * yk
            k += 1
        iters.append(n_done)
        xref = xk
    info = dict(normU=float(normU), normest_iters=int(ne_it), iters=iters, trace_f=[float(v) for v in trace],
                lam=float(lam), mu0=float(mu0))
    return xref, info


CASES = {
    # name: (generator kwargs, nesta kwargs)
    "g1_f3c2_8x8_stop": (dict(H=8, W=8, C=2, F=3, accel=2.0, n_centre=2, slice_id=1),
                         dict(maxIntIter=3, maxIter=30, stoptest=1, miniter=7)),
    "g2_f4c3_12x10_fixed": (dict(H=12, W=10, C=3, F=4, accel=3.0, n_centre=4, slice_id=2),
                            dict(maxIntIter=4, maxIter=10, stoptest=0, miniter=7)),
    "g3_f5c2_16x16_unmasked_y": (dict(H=16, W=16, C=2, F=5, accel=4.0, n_centre=4, slice_id=3),
                                 dict(maxIntIter=2, maxIter=15, stoptest=0, miniter=5)),
    "g4_f6c1_20x14_single_coil": (dict(H=20, W=14, C=1, F=6, accel=2.0, n_centre=4, slice_id=4),
                                  dict(maxIntIter=3, maxIter=12, stoptest=1, miniter=4)),
}


def main():
    index = {}
    for name, (gk, nk) in CASES.items():
        sl = synth.make_cine(**gk)
        y = sl.y
        if "unmasked" in name:
            # k-space with data on unsampled rows too: exercises r = -y outside the mask
            rng = np.random.default_rng(7)
            y = (y + (1 - sl.M[:, None]) * 0.01 * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(
                np.complex64)
        x, info = nesta_numpy(y, sl.S, sl.M, **nk)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), y=y, S=sl.S, M=sl.M, x=x.astype(np.complex128))
        index[name] = dict(gen=gk, nesta=nk, **info)
        print(name, info["iters"], info["normU"])
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(index, fh, indent=1)


if __name__ == "__main__":
    main()
