"""ORACLE — test infrastructure only.

CPU restatement of OpenCLIPER's groupwise B-spline registration (GroupwiseRegistration: the cardiac motion estimator
that fills ArgumentsMotionCompensation for motion-compensated NESTA, tests/MRIRecon.cpp:209-282).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, always as the checker; the library never
calls it.

How it restates the reference
-----------------------------
The host steps run in the reference's order: GroupwiseRegistration::init (GroupwiseRegistration.cpp:93-419),
setMask (:485-535), launch (:550-615), creabspline (groupwiseRegistration/creabspline.cpp:57-258), optimizer
(groupwiseRegistration/optimizer.cpp:58-375), evolution (GroupwiseRegistration.cpp:456-483) and stopCondition
(:537-548).  Every device kernel (src/kernels/optimizer.cl, creabspline.cl, reductions.cl, initialize.cl) is replayed
over its global range with the kernel's own index arithmetic, vectorised over the work items with numpy: the function
for a kernel is named after it, cites its lines, and takes the same buffers (flat arrays) and scalar arguments that
the kernel's launcher (groupwiseRegistration/**/X.cpp launch()) passes, with the global size that launcher uses.  Host
arithmetic that the reference does in float (B-spline tables, coefficient ranges, Wn, H, Tdif) is rounded to float32
where the reference rounds; device work is done in float64.

Differences from the reference, all at places where the reference is undefined:
* costReduction (reductions.cl:144-156) runs the same serial sum in every one of its H*W work items, all adding into
  cost[iter] at once; here the sum is taken once.
* Bilinear neighbours outside the frame (optimizer.cl:236-295, 832-867) are read by the reference out of bounds;
  here they read as zero (as in the library).

Supported configuration (anything else raises NotImplementedError): a square image, a circular ROI (radius > 0),
cubic splines (E = 3), Dp[0] == Dp[1], and a geometry whose support ranges and tap indices stay inside the frame and
inside the control grid.  These are the conditions under which the reference's kernels index consistently: the
wholebox ROI (radius 0) fills bound_box in a different order (GroupwiseRegistration.cpp:121-124) from the one every
consumer reads (:518-521), and E = 2 has a typo in its first derivative (bspline.cpp:186).

Pins (tests/test_gwr_oracle.py)
-------------------------------
* B-spline tables: the reference's cubic uses 0.6666 for 2/3 (bspline.cpp:117), so the taps of a pixel sum to
  1 - (2/3 - 0.6666) * (number of taps with |u - k| < 1), a closed form; first and second derivative taps sum to 0;
  derivative taps match central differences of the value taps.
* T = 0 gives the identity mesh and an interpolated image equal to the input; a constant T shifts every box pixel
  by T times the tap sum.
* The metric gradient (dH with all lambda = 0) matches central finite differences of the cost H(T) on smooth images.
* Registration of frames translated by known sub-pixel shifts recovers the shifts (up to the common mean).
* evolution / stopCondition / sumLambda against hand-computed sequences (sumLambda accumulates into an int).
* The regularisation gradient is constant over frames, so UpdateTransformation's frame-mean projection removes it:
  runs with and without lambda give the same T up to rounding.
Parity unpinned: the index arithmetic of the regularisation terms themselves (Regularization, optimizer.cl:746-831,
and its reductions) beyond the properties above; they reach the result only through SumLambdaMetric, which the
reference runs only when the int-accumulated lambda sum is non-zero.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def _f(v):
    return float(np.float32(v))


# ---------------------------------------------------------------- B-splines (groupwiseRegistration/bspline.cpp)
def bspline3(x):
    """bspline.cpp:113-131: pow() promotes to double in the first branch; the second is float arithmetic."""
    x = np.asarray(x, dtype=np.float32)
    val = np.abs(x)
    v64 = val.astype(np.float64)
    y1 = (0.5 * v64 ** 3 - v64 ** 2 + 0.6666).astype(np.float32)
    t = (np.float32(2) - val)
    y2 = (t * t * t / np.float32(6)).astype(np.float32)
    return np.where(val < 1, y1, np.where(val < 2, y2, np.float32(0))).astype(np.float32)


def b1spline3(x):
    """bspline.cpp:198-219."""
    x = np.asarray(x, dtype=np.float32)
    x64 = x.astype(np.float64)
    a = (1.5 * x64 ** 2 - 2 * x64).astype(np.float32)
    b = (-1.5 * x64 ** 2 - 2 * x64).astype(np.float32)
    c = (-(x - np.float32(2)) * (x - np.float32(2)) / np.float32(2)).astype(np.float32)
    d = ((x + np.float32(2)) * (x + np.float32(2)) / np.float32(2)).astype(np.float32)
    return np.where((x <= 1) & (x >= 0), a,
                    np.where((x >= -1) & (x < 0), b,
                             np.where((x > 1) & (x < 2), c,
                                      np.where((x < -1) & (x > -2), d, np.float32(0))))).astype(np.float32)


def b2spline3(x):
    """bspline.cpp:267-285."""
    x = np.asarray(x, dtype=np.float32)
    val = np.abs(x)
    return np.where(val < 1, np.float32(3) * val - np.float32(2),
                    np.where(val < 2, -val + np.float32(2), np.float32(0))).astype(np.float32)


# ---------------------------------------------------------------- helpers
def _gather(a, idx):
    idx = np.asarray(idx, dtype=np.int64)
    if idx.size and (idx.min() < 0 or idx.max() >= a.size):
        raise NotImplementedError("the reference indexes outside a buffer for this geometry")
    return a[idx]


def _gather_zero(a, idx, ok):
    """Neighbour reads of the interpolators: outside the frame they read zero (the reference is undefined there)."""
    idx = np.where(ok, np.asarray(idx, dtype=np.int64), 0)
    return np.where(ok, a[idx], 0.0)


def _grid(*sizes):
    """Global ids (g0, g1, g2) of an NDRange, flattened, g0 fastest."""
    g = np.meshgrid(*[np.arange(s, dtype=np.int64) for s in sizes], indexing="ij")
    return [x.ravel() for x in g]


def sum_lambda(lam):
    """GroupwiseRegistration.cpp:441-443: std::accumulate(begin, end, 0) keeps an int accumulator."""
    acc = 0
    for v in lam:
        acc = int(acc + np.float32(v))       # int + float -> float, converted back to int (truncation toward 0)
    return acc


def evolution(Tdif, Hdif, Wn, Dif, H, it, flagW):
    """GroupwiseRegistration.cpp:456-483: float sum of squares of Dif, H difference, Wn x1.2 or /2."""
    s = np.float32(0)
    for v in np.asarray(Dif, dtype=np.float32).ravel():
        s = np.float32(s + np.float32(v * v))
    Tdif[it - 1] = np.float32(math.sqrt(s))
    Hdif[it - 1] = np.float32(np.float32(H[it - 1]) - np.float32(H[it]))
    if flagW:
        Wn = np.float32(float(Wn) * 1.2) if Hdif[it - 1] > 0 else np.float32(float(Wn) / 2)
    return Wn


def stop_condition(p1, p3, p2, p4, p5, p6):
    """GroupwiseRegistration.cpp:537-548."""
    return (p5 == p6) or ((p1 < p2) and (p3 < p4))


def set_mask(N, radius):
    """GroupwiseRegistration.cpp:485-535 (square image): circle of the given radius about ((N+1)/2, (N+1)/2)."""
    center = np.float32(N + 1) / np.float32(2)
    i = np.arange(N)[:, None]          # loop over columns (fastest index of X)
    j = np.arange(N)[None, :]
    Cv = ((np.float32(1) - center + j.astype(np.float32)).astype(np.float64) ** 2).astype(np.float32)
    Dv = ((np.float32(1) - center + i.astype(np.float32)).astype(np.float64) ** 2).astype(np.float32)
    inside = (Cv + Dv) <= int(radius ** 2)
    X = np.zeros(N * N, dtype=np.int64)
    ii, jj = np.nonzero(inside)
    X[ii + jj * N] = 1
    bb = [int(ii.min()) + 1, int(jj.min()) + 1, int(ii.max()) + 1, int(jj.max()) + 1]
    return X, bb, int(inside.sum())


class Registration:
    """GroupwiseRegistration with its InitParameters (W, flagW, radius, E, Dp, nmax, et, eh, lambda)."""

    def __init__(self, N, F, W=1.0, flagW=True, radius=40, E=3, Dp=(4, 4), nmax=100, et=0.01, eh=0.005,
                 lam=(0.0, 0.005, 0.0, 0.5)):
        if radius == 0:
            raise NotImplementedError("wholebox ROI: bound_box order differs from its consumers")
        if E != 3:
            raise NotImplementedError("only cubic splines")
        if Dp[0] != Dp[1]:
            raise NotImplementedError("Dp[0] != Dp[1]")
        if F < 2:
            raise NotImplementedError("the frame shifts of the temporal terms need two frames")
        self.N, self.F, self.nt = N, F, 2
        self.E, self.Dp = E, (int(Dp[0]), int(Dp[1]))
        self.nmax, self.et, self.eh = int(nmax), _f(et), _f(eh)
        self.lam = [_f(v) for v in lam]
        self.flagW = bool(flagW)
        # init(): ROI, bound box, sizes (GroupwiseRegistration.cpp:110-176)
        self.X, bb, self.areaX = set_mask(N, radius)
        self.bb = bb
        self.longr1 = bb[2] - bb[0] + 1
        self.longr2 = bb[3] - bb[1] + 1
        self.margin = [N // 40, N // 40, F // 40]
        self.longr1margin = (bb[3] + self.margin[0] + 1) - (bb[1] - self.margin[0] - 1) + 1
        self.longr2margin = (bb[2] + self.margin[1] + 1) - (bb[0] - self.margin[1] - 1) + 1
        self.Wn0 = np.float32(np.float32(W) / np.float32(self.areaX))
        x = np.zeros(2 * N * N, dtype=np.int64)
        i = np.arange(N)[:, None]
        j = np.arange(N)[None, :]
        x[(i + j * N).ravel()] = np.broadcast_to(j + 1, (N, N)).ravel()
        x[(i + j * N + N * N).ravel()] = np.broadcast_to(i + 1, (N, N)).ravel()
        self.x = x
        # Bspline ts (GroupwiseRegistration.cpp:177-196)
        self.c = [np.float32(bb[0] + bb[2]) / np.float32(2), np.float32(bb[1] + bb[3]) / np.float32(2)]
        self.rp = [np.float32((E + 1) * self.Dp[l] // 2) for l in range(2)]
        self.C = [[int(math.floor(np.float32(np.float32(abs(self.c[l] - np.float32(bb[l + 2 * b]))) + self.rp[l])
                                  / np.float32(self.Dp[l]))) for b in range(2)] for l in range(2)]
        self.dp0 = int(2 * math.ceil(self.rp[0]) + 1)
        self.dp1 = int(2 * math.ceil(self.rp[1]) + 1)
        self.Cp1 = self.C[1][0] + self.C[1][1] + 1
        self.Cp0 = self.C[0][0] + self.C[0][1] + 1
        if self.Cp0 != self.Cp1 or self.dp0 != self.dp1 or self.longr1 != self.longr2:
            raise NotImplementedError("asymmetric control grid")
        if bb[1] - self.margin[0] - 2 < 0 or bb[3] + self.margin[0] > N - 1:
            raise NotImplementedError("the interpolation margin leaves the frame")
        self.creabspline()

    # ------------------------------------------------------------ creabspline (creabspline.cpp:57-258)
    def creabspline(self):
        N, E, nt = self.N, self.E, self.nt
        Cp0, Cp1, dp0, dp1 = self.Cp0, self.Cp1, self.dp0, self.dp1
        c, Dp, rp = self.c, self.Dp, self.rp
        # controlPointsLocation (creabspline.cl:50-60), pu dims {Cp0, Cp1, nt}, global (Cp0, Cp1)
        pu = np.zeros(Cp0 * Cp1 * nt, dtype=np.float32)
        gi, gj = _grid(Cp0, Cp1)
        width, height = Cp0, Cp1
        pu[gj + gi * width] = (c[0] + np.float32(Dp[0]) * (-self.C[0][0] + gj).astype(np.float32)).astype(np.float32)
        pu[gj + gi * width + width * height] = (c[1] + np.float32(Dp[1]) * (-self.C[1][0] + gi).astype(np.float32)
                                                ).astype(np.float32)
        self.pu = pu
        # coef, E odd (creabspline.cpp:96-109)
        x = self.x
        coef = np.zeros(N * 2 * nt, dtype=np.float32)
        i = np.arange(N)
        for b in range(2):
            sgn = (-1.0) ** (b + 1)
            add = ((E + sgn) / 2) * sgn
            u0 = ((x[0 + i * N].astype(np.float32) - c[0]) / np.float32(Dp[0])).astype(np.float32)
            u1 = ((x[i + N * N].astype(np.float32) - c[1]) / np.float32(Dp[1])).astype(np.float32)
            coef[i + b * N + 0 * N * 2] = np.floor(u0) + add
            coef[i + b * N + 1 * N * 2] = np.floor(u1) + add
        self.coef_raw = coef
        # BBAux dims {N, E+1, 3, nt} (creabspline.cpp:113-131)
        E1 = E + 1
        BBAux = np.zeros(N * E1 * 3 * nt, dtype=np.float32)
        for l in range(nt):
            for n in range(N):
                lo = int(coef[n + 0 * N + l * N * 2])
                hi = int(coef[n + 1 * N + l * N * 2])
                for k in range(lo, hi + 1):
                    arg = np.float32(np.float32(np.float32(n + 1) - c[l]) / np.float32(Dp[l])) - np.float32(k)
                    base = (k - lo) + n * E1 + l * E1 * N * 3
                    BBAux[base + 0 * E1 * N] = bspline3(arg)
                    BBAux[base + 1 * E1 * N] = b1spline3(arg)
                    BBAux[base + 2 * E1 * N] = b2spline3(arg)
        self.BBAux = BBAux
        # createBB (creabspline.cl:77-98): output dims {E+1, E+1, N, N, (2)}, global (E+1, E+1, N*N), dim2 = N
        dim0 = dim1 = E1
        dim2 = N
        dim3 = N
        gi, gj, g2 = _grid(E1, E1, N * N)
        m = g2 % dim2
        n = g2 // dim2
        idx = gj + gi * dim1 + m * dim1 * dim0 + n * dim1 * dim0 * dim2
        sz = E1 * E1 * N * N
        auxBB = np.zeros(sz, np.float32)
        auxBB1 = np.zeros(2 * sz, np.float32)
        auxBB2 = np.zeros(2 * sz, np.float32)
        auxBB11 = np.zeros(sz, np.float32)

        def BA(ii, coord, t, l):
            return BBAux[ii + coord * dim0 + t * dim0 * dim2 + l * dim0 * dim2 * 3]
        auxBB[idx] = BA(gi, n, 0, 1) * BA(gj, m, 0, 0)
        auxBB1[idx] = BA(gi, n, 0, 1) * BA(gj, m, 1, 0)
        auxBB1[idx + sz] = BA(gi, n, 1, 1) * BA(gj, m, 0, 0)
        auxBB2[idx] = BA(gi, n, 0, 1) * BA(gj, m, 2, 0)
        auxBB2[idx + sz] = BA(gi, n, 2, 1) * BA(gj, m, 0, 0)
        auxBB11[idx] = BA(gi, n, 1, 1) * BA(gj, m, 1, 0)
        del dim3
        # coefbuffer (creabspline.cpp:140-151)
        coefb = np.zeros(N * 2 * nt, dtype=np.float32)
        for l in range(2):
            off = coef[self.bb[l] + 0 * N + l * N * 2]
            for b in range(2):
                coefb[i + b * N + l * N * 2] = coef[i + b * N + l * N * 2] + np.float32(1) - off
        self.coef = coefb
        # permuteBB (creabspline.cl:125-154): dims of BB1 {N, N, F, nt, E+1, E+1, 2}, global (N*N, F*nt, E1*E1)
        F = self.F
        dim0, dim2, dim4 = N, F, E1
        dim1, dim3, dim5 = N, nt, E1
        g0, g1, g2 = _grid(N * N, F * nt, E1 * E1)
        ii = g0 % dim0
        jj = g0 // dim0
        kk = g1 % dim2
        mm = g1 // dim2
        nn = g2 % dim4
        oo = g2 // dim4
        szBB = N * N * F * nt * E1 * E1
        self.BB = np.zeros(szBB, np.float32)
        self.BB1 = np.zeros(2 * szBB, np.float32)
        self.BB2 = np.zeros(2 * szBB, np.float32)
        self.BB11 = np.zeros(szBB, np.float32)
        for l in range(2):
            ip = nn + oo * dim5 + jj * dim5 * dim4 + ii * dim5 * dim4 * dim0 + l * dim5 * dim4 * dim0 * dim1
            ix = (jj + ii * dim1 + kk * dim1 * dim0 + mm * dim1 * dim0 * dim2 + nn * dim1 * dim0 * dim2 * dim3
                  + oo * dim1 * dim0 * dim2 * dim3 * dim4 + l * dim1 * dim0 * dim2 * dim3 * dim4 * dim5)
            self.BB1[ix] = np.float32(Dp[l]) * auxBB1[ip]
            self.BB2[ix] = np.float32(Dp[l] * Dp[l]) * auxBB2[ip]
        ip = nn + oo * dim5 + jj * dim5 * dim4 + ii * dim5 * dim4 * dim0
        ix = jj + ii * dim1 + kk * dim1 * dim0 + mm * dim1 * dim0 * dim2 + nn * dim1 * dim0 * dim2 * dim3 + \
            oo * dim1 * dim0 * dim2 * dim3 * dim4
        self.BB11[ix] = np.float32(Dp[0] * Dp[1]) * auxBB11[ip]
        self.BB[ix] = auxBB[ip]
        # coefgbuffer (creabspline.cpp:161-167), before the clamp at :226-236
        Cg = Cp1
        coefg = np.zeros(Cg * 2 * nt, dtype=np.float32)
        crp0 = np.float32(math.ceil(rp[0]))
        crp1 = np.float32(math.ceil(rp[1]))
        for ii_ in range(Cg):
            for jj_ in range(2):
                coefg[ii_ + jj_ * Cg + 0] = np.float32(np.ceil(np.float32(pu[ii_] - crp0))) + np.float32(jj_ * 2) * crp0
                coefg[ii_ + jj_ * Cg + Cg * 2] = np.float32(np.ceil(np.float32(pu[0 + ii_ * Cp1 + Cp1 * Cp0] - crp1))) + \
                    np.float32(jj_ * 2) * crp1
        xbis2 = np.zeros(Cg * 2, dtype=np.float32)
        xbis2[:Cg] = pu[np.arange(Cg)]
        xbis2[Cg:] = pu[np.arange(Cg) * Cp1 + Cp1 * Cp0]
        # BBAux1 dims {Cp1, dp0, 3, 2} (creabspline.cpp:176-212)
        BBAux1 = np.zeros(Cg * dp0 * 3 * 2, dtype=np.float32)
        for n in range(Cg):
            lo = int(coefg[n])
            hi = int(coefg[n + Cg])
            for k in range(lo, hi + 1):
                base = (k - lo) + n * dp0
                a0 = np.float32(np.float32(np.float32(k) - xbis2[n]) / np.float32(Dp[0]))
                a1 = np.float32(np.float32(np.float32(k) - xbis2[n + Cg]) / np.float32(Dp[1]))
                BBAux1[base + 0 * dp0 * Cg + 0 * dp0 * Cg * 3] = bspline3(a0)
                BBAux1[base + 0 * dp0 * Cg + 1 * dp0 * Cg * 3] = bspline3(a1)
                BBAux1[base + 1 * dp0 * Cg + 0 * dp0 * Cg * 3] = b1spline3(a0)
                BBAux1[base + 1 * dp0 * Cg + 1 * dp0 * Cg * 3] = b1spline3(a1)
                BBAux1[base + 2 * dp0 * Cg + 0 * dp0 * Cg * 3] = b2spline3(a0)
                BBAux1[base + 2 * dp0 * Cg + 1 * dp0 * Cg * 3] = b2spline3(a1)
        self.BBAux1 = BBAux1
        # createBBg (creabspline.cl:174-193): dims of auxBB1g {dp1, dp0, Cp1, Cp0, 2}, global (dp1, dp0, Cp1*Cp0)
        d0, d1, d2, d3 = dp1, dp0, Cp1, Cp0
        gi, gj, g2 = _grid(d0, d1, d2 * d3)
        n = g2 % d2
        m = g2 // d2
        idx = gj + gi * d1 + n * d1 * d0 + m * d1 * d0 * d2
        szg = dp1 * dp0 * Cp1 * Cp0

        def B1(s, ctrl, t, l):
            return BBAux1[s + ctrl * d1 + t * d1 * d2 + l * d1 * d2 * 3]
        self.BBg = np.zeros(szg, np.float32)
        auxBB1g = np.zeros(2 * szg, np.float32)
        auxBB2g = np.zeros(2 * szg, np.float32)
        auxBB11g = np.zeros(szg, np.float32)
        self.BBg[idx] = B1(gi, m, 0, 1) * B1(gj, n, 0, 0)
        auxBB11g[idx] = np.float32(2 * Dp[0] * Dp[1]) * B1(gi, m, 1, 1) * B1(gj, n, 1, 0)
        auxBB1g[idx] = np.float32(2 * Dp[0]) * B1(gi, m, 0, 1) * B1(gj, n, 1, 0)
        auxBB1g[idx + szg] = np.float32(2 * Dp[1]) * B1(gi, m, 1, 1) * B1(gj, n, 0, 0)
        auxBB2g[idx] = np.float32(2 * Dp[0] * Dp[0]) * B1(gi, m, 0, 1) * B1(gj, n, 2, 0)
        auxBB2g[idx + szg] = np.float32(2 * Dp[1] * Dp[1]) * B1(gi, m, 2, 1) * B1(gj, n, 0, 0)
        # repmatBBg (creabspline.cl:217-245): dims of aux1BB1g {dp1, dp0, Cp1, Cp0, 2, F, 2}
        dim0, dim2, dim4, dim6 = dp1, Cp1, 2, 2
        dim1, dim3, dim5 = dp0, Cp0, F
        g0, g1, g2 = _grid(dp1 * dp0, Cp1 * Cp0, 2 * F)
        ii, jj = g0 % dim0, g0 // dim0
        kk, ll = g1 % dim2, g1 // dim2
        mm, nn = g2 % dim4, g2 // dim4
        s7 = dp1 * dp0 * Cp1 * Cp0 * 2 * F * 2
        aux1BB1g = np.zeros(s7, np.float32)
        aux1BB2g = np.zeros(s7, np.float32)
        aux1BB11g = np.zeros(dp1 * dp0 * Cp1 * Cp0 * F * 2, np.float32)
        for o in range(dim6):
            ia = jj + ii * dim1 + kk * dim0 * dim1 + ll * dim0 * dim1 * dim2 + mm * dim0 * dim1 * dim2 * dim3
            ix = ia + nn * dim0 * dim1 * dim2 * dim3 * dim4 + o * dim0 * dim1 * dim2 * dim3 * dim4 * dim5
            aux1BB1g[ix] = auxBB1g[ia]
            aux1BB2g[ix] = auxBB2g[ia]
            ia = jj + ii * dim1 + kk * dim0 * dim1 + ll * dim0 * dim1 * dim2
            ix = ia + nn * dim0 * dim1 * dim2 * dim3 + o * dim0 * dim1 * dim2 * dim3 * dim5
            aux1BB11g[ix] = auxBB11g[ia]
        # permuteBBg (creabspline.cl:267-293): dims of BB1g {dp1, dp0, F, Cp1, Cp0, 2, 2}
        dim0, dim2, dim4, dim6 = dp1, F, Cp0, 2
        dim1, dim3, dim5 = dp0, Cp1, 2
        g0, g1, g2 = _grid(dp1 * dp0, F * Cp1, Cp0 * 2)
        ii, jj = g0 % dim0, g0 // dim0
        kk, ll = g1 % dim2, g1 // dim2
        mm, nn = g2 % dim4, g2 // dim4
        self.BB1g = np.zeros(s7, np.float32)
        self.BB2g = np.zeros(s7, np.float32)
        self.BB11g = np.zeros(dp1 * dp0 * F * Cp1 * Cp0 * 2, np.float32)
        for o in range(dim6):
            ia = (jj + ii * dim1 + ll * dim0 * dim1 + mm * dim0 * dim1 * dim3 + o * dim0 * dim1 * dim3 * dim4
                  + kk * dim0 * dim1 * dim3 * dim4 * dim6 + nn * dim0 * dim1 * dim3 * dim4 * dim6 * dim2)
            ix = (jj + ii * dim1 + kk * dim0 * dim1 + ll * dim0 * dim1 * dim2 + mm * dim0 * dim1 * dim2 * dim3
                  + nn * dim0 * dim1 * dim2 * dim3 * dim4 + o * dim0 * dim1 * dim2 * dim3 * dim4 * dim5)
            self.BB1g[ix] = aux1BB1g[ia]
            self.BB2g[ix] = aux1BB2g[ia]
        ia = (jj + ii * dim1 + ll * dim0 * dim1 + mm * dim0 * dim1 * dim3 + kk * dim0 * dim1 * dim3 * dim4
              + nn * dim0 * dim1 * dim3 * dim4 * dim2)
        ix = jj + ii * dim1 + kk * dim0 * dim1 + ll * dim0 * dim1 * dim2 + mm * dim0 * dim1 * dim2 * dim3 + \
            nn * dim0 * dim1 * dim2 * dim3 * dim4
        self.BB11g[ix] = aux1BB11g[ia]
        # clamp coefg to [1, N] (creabspline.cpp:226-236)
        self.coefg_raw = coefg.copy()
        coefg = np.clip(coefg, 1, N).astype(np.float32)
        if not np.array_equal(coefg, self.coefg_raw) or (self.coefg_raw[:Cg] - 1 + dp0 > N).any() or \
                (self.coefg_raw[Cg * 2:Cg * 3] - 1 + dp0 > N).any():
            raise NotImplementedError("a control point's support leaves the frame")
        self.coefg = coefg

    # ------------------------------------------------------------ kernels of the optimizer loop
    def r1(self):
        return np.arange(self.longr1) + (self.bb[0] - 1)

    def r2(self):
        return np.arange(self.longr2) + (self.bb[1] - 1)

    def r1margin(self):
        return (self.bb[1] - self.margin[0] - 1) + np.arange(self.longr1margin) - 1

    def r2margin(self):
        return (self.bb[0] - self.margin[1] - 1) + np.arange(self.longr2margin) - 1

    def transformationAux(self, T):
        """optimizer.cl:565-606; TransformationAux.cpp launch: TAux dims {longr2, longr1, F, nt, E+1, E+1},
        global (longr2*longr1, F*nt, (E+1)^2), args (dim0 longr2, dim2 F, dim4 E+1, tam0 N, C1 Cp1, C0 Cp0)."""
        E1, F, nt, N = self.E + 1, self.F, self.nt, self.N
        dim0, dim2, dim4 = self.longr2, F, E1
        dim1, dim3 = self.longr1, nt
        tam0, C1, C0 = N, self.Cp1, self.Cp0
        g0, g1, g2 = _grid(dim0 * dim1, dim2 * dim3, dim4 * E1)
        i, j = g0 % dim0, g0 // dim0
        m, n = g1 % dim2, g1 // dim2
        l, k = g2 % dim4, g2 // dim4
        r1, r2 = self.r1(), self.r2()
        idxTAux = j + i * dim1 + m * dim0 * dim1 + n * dim0 * dim1 * dim2 + l * dim0 * dim1 * dim2 * dim3 + \
            k * dim0 * dim1 * dim2 * dim3 * dim4
        cf = self.coef
        idxT = (l + cf[r1[j] + 0 * tam0 + 0 * tam0 * 2].astype(np.int64) - 1) + \
            (k + cf[r2[i] + 0 * tam0 + 1 * tam0 * 2].astype(np.int64) - 1) * C0 + m * C1 * C0 + n * C1 * C0 * dim2
        TAux = np.zeros(dim0 * dim1 * F * nt * E1 * E1)
        TAux[idxTAux] = _gather(T, idxT)
        return TAux

    def transformation(self, TAux):
        """optimizer.cl:608-646 (Deformation.cpp launch): global (longr2, longr1, F*nt) over TAux, BB dims
        {N, N, F, nt, E+1, E+1}; xn += TAux * BB over the (E+1)^2 taps.  Followed by transformationAdjust
        (optimizer.cl:648-671): xn += the mesh."""
        N, F, nt, E1 = self.N, self.F, self.nt, self.E + 1
        dT0, dT1, dT2, dT3 = self.longr2, self.longr1, F, nt
        dimBB0, dimBB1 = N, N
        i, j, g2 = _grid(dT0, dT1, dT2 * dT3)
        k, l = g2 % dT2, g2 // dT2
        r1, r2 = self.r1(), self.r2()
        idx = r1[j] + r2[i] * dimBB1 + k * dimBB0 * dimBB1 + l * dimBB0 * dimBB1 * dT2
        xn = np.zeros(N * N * F * nt)
        acc = np.zeros(idx.size)
        for m in range(E1):
            for n in range(E1):
                iT = j + i * dT1 + k * dT0 * dT1 + l * dT0 * dT1 * dT2 + m * dT0 * dT1 * dT2 * dT3 + \
                    n * dT0 * dT1 * dT2 * dT3 * E1
                iB = r1[j] + r2[i] * dimBB1 + k * dimBB0 * dimBB1 + l * dimBB0 * dimBB1 * dT2 + \
                    m * dimBB0 * dimBB1 * dT2 * nt + n * dimBB0 * dimBB1 * dT2 * nt * E1
                acc += TAux[iT] * self.BB[iB].astype(np.float64)
        xn[idx] = acc
        # transformationAdjust, global (N, N, F*nt): xn[j + i*N + ...] += x[i + j*N + l*N*N]
        i, j, g2 = _grid(N, N, F * nt)
        k, l = g2 % F, g2 // F
        xn[j + i * N + k * N * N + l * N * N * F] += self.x[i + j * N + l * N * N]
        return xn

    def deform(self, T):
        return self.transformation(self.transformationAux(T))

    def interpolator(self, I0, xn):
        """optimizer.cl:832-867 after copyDataGPU (initialize.cl:76-80): box + margin pixels of every frame
        resampled bilinearly at xn.  I0: real image [F*N*N] (the imaginary part of the normalised data is 0)."""
        N, F = self.N, self.F
        tam0 = tam1 = N
        out = I0.astype(np.float64).copy()
        i, j, k = _grid(self.longr1margin, self.longr2margin, F)
        r1m, r2m = self.r1margin(), self.r2margin()
        hor = self.x[r1m[i] + r2m[j] * tam0] - 1
        ver = self.x[r1m[i] + r2m[j] * tam0 + tam0 * tam1] - 1
        nhor = xn[r2m[j] + r1m[i] * tam1 + k * tam0 * tam1] - 1
        nver = xn[r2m[j] + r1m[i] * tam1 + k * tam0 * tam1 + tam0 * tam1 * F] - 1
        out[hor + ver * tam1 + k * tam0 * tam1] = self._bilinear(I0, nhor, nver, k)
        return out

    def _bilinear(self, img, nhor, nver, k):
        N = self.N
        ph = np.floor(nhor).astype(np.int64)
        pv = np.floor(nver).astype(np.int64)
        qh, qv = ph + 1, pv + 1

        def px(h, v):
            ok = (h >= 0) & (h < N) & (v >= 0) & (v < N)
            return _gather_zero(img, h + v * N + k * N * N, ok)
        a1 = (qh - nhor) * px(ph, pv) + (nhor - ph) * px(qh, pv)
        a2 = (qh - nhor) * px(ph, qv) + (nhor - ph) * px(qh, qv)
        return (qv - nver) * a1 + (nver - pv) * a2

    def metric(self, I):
        """optimizer.cl:99-118: per-pixel variance over frames (mean3D, optimizer.cl:39-50)."""
        N, F = self.N, self.F
        v = I.reshape(F, N * N)
        mean = v.sum(axis=0) / F
        return ((v - mean) ** 2).sum(axis=0) / F

    def gradient(self, I):
        """optimizer.cl:189-234 (Gradient.cpp, global (W, H, F)): central differences, one-sided at the edges;
        d1 along columns, d2 along rows."""
        N, F = self.N, self.F
        v = I.reshape(F, N, N)
        d1 = np.empty_like(v)
        d2 = np.empty_like(v)
        d1[:, :, 1:-1] = (v[:, :, 2:] - v[:, :, :-2]) / 2
        d1[:, :, 0] = v[:, :, 1] - v[:, :, 0]
        d1[:, :, -1] = v[:, :, -1] - v[:, :, -2]
        d2[:, 1:-1, :] = (v[:, 2:, :] - v[:, :-2, :]) / 2
        d2[:, 0, :] = v[:, 1, :] - v[:, 0, :]
        d2[:, -1, :] = v[:, -1, :] - v[:, -2, :]
        return d1.ravel(), d2.ravel()

    def gradientInterpolator(self, d1, d2, xn):
        """Initdx (initialize.cl:104-119) then gradientInterpolator (optimizer.cl:236-295): dx = [d1, d2], box +
        margin pixels resampled at xn."""
        N, F = self.N, self.F
        tam0 = tam1 = N
        dx = np.concatenate([d1, d2]).astype(np.float64)
        i, j, k = _grid(self.longr1margin, self.longr2margin, F)
        r1m, r2m = self.r1margin(), self.r2margin()
        hor = self.x[r1m[i] + r2m[j] * tam0] - 1
        ver = self.x[r1m[i] + r2m[j] * tam0 + tam0 * tam1] - 1
        nhor = xn[r2m[j] + r1m[i] * tam1 + k * tam0 * tam1] - 1
        nver = xn[r2m[j] + r1m[i] * tam1 + k * tam0 * tam1 + tam0 * tam1 * F] - 1
        o = hor + ver * tam1 + k * tam0 * tam1
        dx[o] = self._bilinear(d1, nhor, nver, k)
        dx[o + N * N * F] = self._bilinear(d2, nhor, nver, k)
        return dx

    def gradientMetric(self, I):
        """optimizer.cl:297-315: dy = (I - mean over frames) * 2/F."""
        N, F = self.N, self.F
        v = I.reshape(F, N * N)
        factor = np.float32(2.0 / F)
        return ((v - v.sum(axis=0) / F) * float(factor)).ravel()

    def permuteTAux(self, TAux):
        """optimizer.cl:673-704 (PermuteTAux.cpp: global (longr1*longr2, F*nt, (E+1)^2), dim0 = longr1)."""
        F, nt, E1 = self.F, self.nt, self.E + 1
        dim0, dim2, dim4 = self.longr1, F, E1
        dim1, dim3 = self.longr2, nt
        g0, g1, g2 = _grid(dim0 * dim1, dim2 * dim3, dim4 * E1)
        i, j = g0 % dim0, g0 // dim0
        k, l = g1 % dim2, g1 // dim2
        m, n = g2 % dim4, g2 // dim4
        rest = k * dim0 * dim1 + l * dim0 * dim1 * dim2 + m * dim0 * dim1 * dim2 * dim3 + \
            n * dim0 * dim1 * dim2 * dim3 * dim4
        out = np.zeros_like(TAux)
        out[j + i * dim1 + rest] = TAux[i + j * dim0 + rest]
        return out

    def shiftTAux(self, P):
        """optimizer.cl:706-744: copies of TAuxPerm shifted by one and two frames, circularly."""
        F, nt, E1 = self.F, self.nt, self.E + 1
        dim0, dim2, dim4 = self.longr1, F, E1
        dim1, dim3 = self.longr2, nt
        g0, g1, g2 = _grid(dim0 * dim1, dim2 * dim3, dim4 * E1)
        i, j = g0 % dim0, g0 // dim0
        k, l = g1 % dim2, g1 // dim2
        m, n = g2 % dim4, g2 // dim4
        base = j + i * dim1 + l * dim0 * dim1 * dim2 + m * dim0 * dim1 * dim2 * dim3 + \
            n * dim0 * dim1 * dim2 * dim3 * dim4
        s1 = np.where(k == 0, dim2 - 1, k - 1)
        s2 = np.where(k == 0, dim2 - 2, np.where(k == 1, dim2 - 1, k - 2))
        idx = base + k * dim0 * dim1
        S1 = np.zeros_like(P)
        S2 = np.zeros_like(P)
        S1[idx] = _gather(P, base + s1 * dim0 * dim1)
        S2[idx] = _gather(P, base + s2 * dim0 * dim1)
        return S1, S2

    def regularization(self, P, S1, S2):
        """optimizer.cl:746-830 (Regularization.cpp: global (longr1, longr2, F*nt), dtau dims {N, N, F, nt(, 2)}).
        Returns dtaux, dtaux2 (2*N*N*F*nt), dtauxy, dtaut, dtaut2 (N*N*F*nt)."""
        N, F, nt, E1 = self.N, self.F, self.nt, self.E + 1
        d0 = d1 = N
        d2, d3 = F, nt
        dT0, dT1 = self.longr1, self.longr2
        i, j, g2 = _grid(dT0, dT1, d2 * d3)
        k, l = g2 % d2, g2 // d2
        r1, r2 = self.r1(), self.r2()
        n4 = N * N * F * nt
        dtaux = np.zeros(2 * n4)
        dtaux2 = np.zeros(2 * n4)
        dtauxy = np.zeros(n4)
        dtaut = np.zeros(n4)
        dtaut2 = np.zeros(n4)
        pix = r2[j] + r1[i] * d1 + k * d0 * d1 + l * d0 * d1 * d2

        def iTA(m, n):
            return j + i * dT1 + k * dT0 * dT1 + l * dT0 * dT1 * d2 + m * dT0 * dT1 * d2 * d3 + \
                n * dT0 * dT1 * d2 * d3 * E1

        def iB(m, n):
            return pix + m * d0 * d1 * d2 * d3 + n * d0 * d1 * d2 * d3 * E1
        for p in range(2):
            idx = pix + p * n4
            a1 = np.zeros(idx.size)
            a2 = np.zeros(idx.size)
            for m in range(E1):
                for n in range(E1):
                    t = P[iTA(m, n)]
                    b = iB(m, n) + p * n4 * E1 * E1
                    a1 += t * self.BB1[b]
                    a2 += t * self.BB2[b]
            dtaux[idx] += a1
            dtaux2[idx] += a2
            sel = l == p
            idd = (r2[j] + r1[i] * d1 + k * d0 * d1 + p * d0 * d1 * d2 + p * d0 * d1 * d2 * d3)[sel]
            dtaux[idd] += 1
        axy = np.zeros(pix.size)
        at = np.zeros(pix.size)
        at2 = np.zeros(pix.size)
        for m in range(E1):
            for n in range(E1):
                ia = iTA(m, n)
                b = iB(m, n)
                axy += P[ia] * self.BB11[b]
                at += (P[ia] - S1[ia]) * self.BB[b]
                at2 += (P[ia] - 2.0 * S1[ia] + S2[ia]) * self.BB[b]
        dtauxy[pix] += axy
        dtaut[pix] += at
        dtaut2[pix] += at2
        return dtaux, dtaux2, dtauxy, dtaut, dtaut2

    def sumLambdaMetric(self, V, dtau):
        """optimizer.cl:138-187 (SumLambdaMetric.cpp: global (longr1, longr2))."""
        dtaux, dtaux2, dtauxy, dtaut, dtaut2 = dtau
        N, F, nt = self.N, self.F, self.nt
        lam = self.lam
        i, j = _grid(self.longr1, self.longr2)
        r1, r2 = self.r1(), self.r2()
        idx2 = r2[j] + r1[i] * N
        s = [np.zeros(idx2.size) for _ in range(5)]
        for k in range(F):
            for l in range(nt):
                idx4 = idx2 + k * N * N + l * N * N * F
                s[2] += lam[1] * 2 * dtauxy[idx4] ** 2
                s[3] += lam[3] * dtaut2[idx4] ** 2
                s[4] += lam[2] * dtaut[idx4] ** 2
                for m in range(2):
                    idx5 = idx4 + m * N * N * F * nt
                    s[0] += lam[0] * dtaux[idx5] ** 2
                    s[1] += lam[1] * dtaux2[idx5] ** 2
        V = V.copy()
        V[idx2] += s[0] + s[1] + s[2] + s[3] + s[4]
        return V

    def cost(self, V):
        """costReduction (reductions.cl:144-156): sum of V over the ROI mask."""
        return float(np.dot(V, self.X))

    def gradientRegularization(self, dtau):
        """optimizer.cl:317-406 (GradientRegularization.cpp: global (dp1*dp0, F*Cp1, Cp0*nt), args dim0 dp1,
        dim2 F, dim4 Cp0, dimdtau0/1 N)."""
        dtaux, dtaux2, dtauxy, dtaut, dtaut2 = dtau
        N, F, nt = self.N, self.F, self.nt
        dim0, dim2, dim4 = self.dp1, F, self.Cp0
        dim1, dim3, dim5 = self.dp0, self.Cp1, nt
        D0 = D1 = N
        D2 = F
        g0, g1, g2 = _grid(dim0 * dim1, dim2 * dim3, dim4 * dim5)
        i, j = g0 % dim0, g0 // dim0
        n, k = g1 % dim2, g1 // dim2
        l, o = g2 % dim4, g2 // dim4
        cg = self.coefg
        range1 = cg[k + 0 * dim3 + 0 * dim3 * 2].astype(np.int64) - 1
        range2 = cg[l + 0 * dim3 + 1 * dim3 * 2].astype(np.int64) - 1
        s6 = dim0 * dim1 * dim2 * dim3 * dim4 * dim5
        out_x = np.zeros(2 * s6)
        out_x2 = np.zeros(2 * s6)
        base = j + i * dim1 + n * dim0 * dim1 + k * dim0 * dim1 * dim2 + l * dim0 * dim1 * dim2 * dim3 + \
            o * dim0 * dim1 * dim2 * dim3 * dim4
        for m in range(2):
            idx = base + m * s6
            idt = (range2 + i) + (range1 + j) * D0 + n * D0 * D1 + o * D0 * D1 * D2 + m * D0 * D1 * D2 * 2
            out_x[idx] = _gather(dtaux, idt) * self.BB1g[idx]
            out_x2[idx] = _gather(dtaux2, idt) * self.BB2g[idx]
        idt = (range2 + i) + (range1 + j) * D0 + n * D0 * D1 + o * D0 * D1 * D2
        out_xy = _gather(dtauxy, idt) * self.BB11g[base]

        def tau(arr, f):
            return _gather(arr, range1 + range2 * D1 + f * D0 * D1 + o * D0 * D1 * D2)
        ijk = j + i * dim1 + k * dim0 * dim1 + l * dim0 * dim1 * dim3
        jik = i + j * dim1 + k * dim0 * dim1 + l * dim0 * dim1 * dim3
        bg_ij = self.BBg[ijk]
        bg_ji = self.BBg[jik]
        n1 = np.where(n + 1 >= dim2, n + 1 - dim2, n + 1)
        n2 = np.where(n + 2 >= dim2, n + 2 - dim2, n + 2)
        last2 = n == dim2 - 2
        last1 = (n == dim2 - 1) & ~last2
        dt = 2 * (tau(dtaut, n) - tau(dtaut, n1)) * np.where(last2, bg_ij, bg_ji)
        # dthetat2: the second difference is circular in every branch
        dt2 = 2 * (tau(dtaut2, n) - 2 * tau(dtaut2, n1) + tau(dtaut2, n2)) * bg_ij
        del last1
        return out_x, out_x2, out_xy, dt, dt2, s6

    def reductions(self, dtheta):
        """reduction7Dto6D (reductions.cl:56-89), reductionsum over dim 5 and then dim 2 (reductions.cl:91-111,
        ReductionSum.cpp: output[i] = sum_k input[i + k*stride(dim)] over the output's element count)."""
        lam = self.lam
        tx, tx2, txy, tt, tt2, s6 = dtheta
        aux = lam[0] * (tx[:s6] + tx[s6:]) + lam[1] * (tx2[:s6] + tx2[s6:]) + lam[1] * 2 * txy + lam[2] * tt + \
            lam[3] * tt2
        dp1, dp0, F, Cp1, Cp0, nt = self.dp1, self.dp0, self.F, self.Cp1, self.Cp0, self.nt
        s5 = dp1 * dp0 * F * Cp1 * Cp0
        aux1 = sum(aux[k * s5: (k + 1) * s5] for k in range(nt))
        nout = dp1 * dp0 * Cp1 * Cp0
        stride2 = dp1 * dp0
        ii = np.arange(nout)
        aux2 = np.zeros(nout)
        for k in range(F):
            aux2 += _gather(aux1, ii + k * stride2)
        return aux2

    def gradient_joint(self, dy, dx, aux2):
        """gradientJointAux (optimizer.cl:408-440: Aux = aux2 over frames) and gradientJoint (optimizer.cl:442-486,
        GradientWithSmoothTerms.cpp: global (dp1*dp0, F*Cp1, Cp0*nt), args dim0 dp1, dim2 F, dim4 Cp0, tam0/1 N)."""
        N, F, nt = self.N, self.F, self.nt
        dim0, dim2, dim4 = self.dp1, F, self.Cp0
        dim1, dim3, dim5 = self.dp0, self.Cp1, nt
        tam0 = tam1 = N
        tam2 = F
        g0, g1, g2 = _grid(dim0 * dim1, dim2 * dim3, dim4 * dim5)
        i, j = g0 % dim0, g0 // dim0
        n, k = g1 % dim2, g1 // dim2
        l, m = g2 % dim4, g2 // dim4
        cg = self.coefg
        range1 = cg[k].astype(np.int64) - 1 + j
        range2 = cg[l + dim3 * 2].astype(np.int64) - 1 + i
        idxdy = range1 + range2 * tam1 + n * tam0 * tam1
        idxdV = j + i * dim1 + n * dim0 * dim1 + k * dim0 * dim1 * dim2 + l * dim0 * dim1 * dim2 * dim3 + \
            m * dim0 * dim1 * dim2 * dim3 * dim4
        idxdx = range1 + range2 * tam1 + n * tam0 * tam1 + m * tam0 * tam1 * tam2
        # Aux[j + i*dim1 + n*dim0*dim1 + k*... + l*...] = aux2[j + i*dim1 + k*dim0*dim1 + l*dim0*dim1*dim3]
        aux = aux2[j + i * dim1 + k * dim0 * dim1 + l * dim0 * dim1 * dim3]
        dtheta = self.BBg[j + i * dim1 + l * dim0 * dim1 + m * dim0 * dim1 * dim3]
        dV = np.zeros(dim0 * dim1 * dim2 * dim3 * dim4 * dim5)
        dV[idxdV] = _gather(dy, idxdy) * _gather(dx, idxdx) * dtheta + aux
        return dV

    def cost6D(self, dV):
        """permutedV (optimizer.cl:488-517), then for every (k, l) of the host loop (optimizer.cpp:296-315):
        auxiliarMask (optimizer.cl:938-964), auxiliarCost (:966-1000), cost6DReduction (reductions.cl:113-142),
        cost6D (optimizer.cl:1002-1025).  All (k, l) are done at once.  Returns dH dims {Cp0, Cp1, F, nt}."""
        dp1, dp0, F, Cp1, Cp0, nt, N = self.dp1, self.dp0, self.F, self.Cp1, self.Cp0, self.nt, self.N
        # permutedV: global (dp1*dp0, F*Cp1, Cp0*nt), dim0 dp1, dim2 F, dim4 Cp0
        g0, g1, g2 = _grid(dp1 * dp0, F * Cp1, Cp0 * nt)
        i, j = g0 % dp1, g0 // dp1
        k, l = g1 % F, g1 // F
        m, n = g2 % Cp0, g2 // Cp0
        idxIn = j + i * dp0 + k * dp1 * dp0 + l * dp1 * dp0 * F + m * dp1 * dp0 * F * Cp1 + n * dp1 * dp0 * F * Cp1 * Cp0
        idxOut = j + i * dp0 + l * dp1 * dp0 + m * dp1 * dp0 * Cp1 + k * dp1 * dp0 * Cp1 * Cp0 + \
            n * dp1 * dp0 * Cp1 * Cp0 * F
        dVp = np.zeros_like(dV)
        dVp[idxOut] = dV[idxIn]
        # host loop k < Cp1 (dV dim 3), l < Cp0 (dV dim 4); work items (i, j, m, n) with global (dp1*dp0, F, nt)
        hk, hl = np.meshgrid(np.arange(Cp1), np.arange(Cp0), indexing="ij")
        hk = hk.ravel()[:, None]
        hl = hl.ravel()[:, None]
        g0, m, n = _grid(dp1 * dp0, F, nt)
        g0, m, n = g0[None, :], m[None, :], n[None, :]
        i, j = g0 % dp1, g0 // dp1
        cg = self.coefg
        idxX = (i + cg[hk + Cp1 * 2].astype(np.int64) - 1) + (j + cg[hl].astype(np.int64) - 1) * N
        Xaux = np.zeros((hk.shape[0], dp1 * dp0 * F * nt))
        np.put_along_axis(Xaux, np.broadcast_to(j + i * dp1 + m * dp1 * dp0 + n * dp1 * dp0 * F, (hk.shape[0], g0.size)),
                          _gather(self.X, idxX).astype(np.float64), axis=1)
        idxA = j + i * dp0 + m * dp1 * dp0 + n * dp1 * dp0 * F
        idxV = j + i * dp0 + hl * dp1 * dp0 + hk * dp1 * dp0 * Cp1 + m * dp1 * dp0 * Cp1 * Cp0 + \
            n * dp1 * dp0 * Cp1 * Cp0 * F
        cost1 = np.zeros_like(Xaux)
        np.put_along_axis(cost1, np.broadcast_to(j + i * dp0 + m * dp1 * dp0 + n * dp1 * dp0 * F, cost1.shape),
                          dVp[idxV] * np.take_along_axis(Xaux, np.broadcast_to(idxA, cost1.shape), axis=1), axis=1)
        # cost6DReduction: cost2[m + i*F + n*dp1*F] = sum_j cost1[i + j*dp0 + m*dp1*dp0 + n*dp1*dp0*F]
        c1 = cost1.reshape(-1, nt, F, dp0, dp1)          # [kl][n][m][j][i]
        cost2 = c1.sum(axis=3)                           # [kl][n][m][i]
        # cost6D: dH[l + k*Cp1 + m*Cp0*Cp1 + n*Cp0*Cp1*F] = sum_i cost2
        s = cost2.sum(axis=3)                            # [kl][n][m]
        dH = np.zeros(Cp0 * Cp1 * F * nt)
        kk = hk.ravel()
        ll = hl.ravel()
        for nn_ in range(nt):
            for mm in range(F):
                dH[ll + kk * Cp1 + mm * Cp0 * Cp1 + nn_ * Cp0 * Cp1 * F] = s[:, nn_, mm]
        return dH

    def updateTransformation(self, T, dH, Wn):
        """optimizer.cl:519-535 (mean4D, optimizer.cl:52-66): Dif = Wn (dH - mean over frames), T -= Dif."""
        Cp0, Cp1, F, nt = self.Cp0, self.Cp1, self.F, self.nt
        d = dH.reshape(nt, F, Cp0 * Cp1)
        Dif = (float(Wn) * (d - d.sum(axis=1, keepdims=True) / F)).ravel()
        return T - Dif, Dif

    # ------------------------------------------------------------ evaluation helpers (one step of the loop)
    def normalize(self, x):
        """launch (GroupwiseRegistration.cpp:566-596) + dataNormalization (initialize.cl:87-102): 500 |x| / max|x|."""
        a = np.abs(np.asarray(x, dtype=np.complex64)).astype(np.float32)
        oldmax = a.max()
        factor = np.float32(np.float32(500) / oldmax)
        return (factor * a).astype(np.float32).ravel().astype(np.float64)

    def cost_at(self, I0, T, dtau=None):
        xn = self.deform(T)
        It = self.interpolator(I0, xn)
        V = self.metric(It)
        if dtau is not None:
            V = self.sumLambdaMetric(V, dtau)
        return self.cost(V), It, xn

    def metric_gradient(self, I0, T, It=None):
        """dH of the cost with the warped image It (default: I0 resampled at T) and no regularisation term."""
        xn = self.deform(T)
        if It is None:
            It = self.interpolator(I0, xn)
        d1, d2 = self.gradient(I0)
        dx = self.gradientInterpolator(d1, d2, xn)
        dy = self.gradientMetric(It)
        aux2 = np.zeros(self.dp1 * self.dp0 * self.Cp1 * self.Cp0)
        return self.cost6D(self.gradient_joint(dy, dx, aux2))

    # ------------------------------------------------------------ launch (optimizer.cpp:58-375)
    def run(self, x, T=None, nmax=None):
        """Registration of the complex frames x [F][N][N].  Returns a dict with T, xn, H, Wn, Tdif, Hdif, iters,
        transformed (the registered normalised frames) and the mesh / bound box."""
        N, F, nt = self.N, self.F, self.nt
        nmax = self.nmax if nmax is None else int(nmax)
        I0 = self.normalize(x)
        if T is None:
            T = np.zeros(self.Cp1 * self.Cp0 * F * nt)
        T = np.asarray(T, dtype=np.float64).copy()
        H = np.zeros(nmax + 1)
        Hdif = np.zeros(nmax, np.float32)
        Tdif = np.zeros(nmax, np.float32)
        Wn = self.Wn0
        Wn_trace = []
        It = I0.copy()                                   # InitIT: transformedData = normInputData
        n4 = N * N * F * nt
        dtau = (np.zeros(2 * n4), np.zeros(2 * n4), np.zeros(n4), np.zeros(n4), np.zeros(n4))
        H[0] = np.float32(self.cost(self.metric(It)))
        d1, d2 = self.gradient(I0)
        p2 = np.float32(np.float32(self.et) * np.float32(F * nt))
        p4 = np.float32(np.float32(H[0]) * np.float32(self.eh))
        use_lambda = sum_lambda(self.lam) != 0
        it = 1
        xn = None
        while True:
            xn = self.deform(T)
            dx = self.gradientInterpolator(d1, d2, xn)
            dy = self.gradientMetric(It)
            aux2 = self.reductions(self.gradientRegularization(dtau))
            dH = self.cost6D(self.gradient_joint(dy, dx, aux2))
            T, Dif = self.updateTransformation(T, dH, Wn)
            TAux = self.transformationAux(T)
            xn = self.transformation(TAux)
            P = self.permuteTAux(TAux)
            S1, S2 = self.shiftTAux(P)
            dtau = self.regularization(P, S1, S2)
            It = self.interpolator(I0, xn)
            V = self.metric(It)
            if use_lambda:
                V = self.sumLambdaMetric(V, dtau)
            H[it] = np.float32(self.cost(V))
            Wn = evolution(Tdif, Hdif, Wn, Dif, H, it, self.flagW)
            Wn_trace.append(float(Wn))
            stop = stop_condition(Tdif[it - 1], Hdif[it - 1], p2, p4, nmax, it)
            if stop:
                break
            it += 1
        return {"T": T, "xn": xn, "H": H[: it + 1], "Tdif": Tdif[:it], "Hdif": Hdif[:it], "Wn": Wn_trace,
                "iters": it, "transformed": It, "x": self.x, "bound_box": list(self.bb), "I0": I0,
                "Cp": self.Cp0}
