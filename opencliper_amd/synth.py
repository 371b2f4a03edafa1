"""Seeded synthetic MRI inputs for the NESTA-TV path (host side, numpy).

Every array is produced on the host from a per-slice seed, so the GPU path and the CPU checkers see the same bytes.
Layouts follow OpenCLIPER's Data classes with W as the fastest dimension:

* image / XData ............ [F][H][W] complex64
* k-space / KData .......... [F][C][H][W] complex64   (coil between frame and pixel, as KData orders its NDArrays)
* sensitivity maps ......... [C][H][W] complex64
* sampling masks ........... [F][H][W] uint8, DC at row/column 0 (no fftshift, the clFFT convention of
  src/processes/FFT.cpp:336-449)

Generators (image intensities U(0.2,1), 10 ellipses, smooth linear phase, Gaussian coil profiles normalised to a
unit root-sum-of-squares, variable-density ky-line masks with 16 fully-sampled centre lines redrawn per frame,
complex Gaussian noise at 1e-3 of the peak k-space magnitude) are this project's own; the reference ships no input
data with its checkout.
"""
from __future__ import annotations

import dataclasses
import numpy as np

BASE_SEED = 20261203
_MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """One step of splitmix64 (Steele, Lea & Flood 2014), used to derive independent per-slice seeds."""
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def slice_rng(base_seed: int, slice_id: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(splitmix64((base_seed ^ slice_id) & _MASK64)))


def _grid(H: int, W: int):
    v = (np.arange(H, dtype=np.float64) - H / 2) / (H / 2)   # rows, [-1, 1)
    u = (np.arange(W, dtype=np.float64) - W / 2) / (W / 2)   # columns, [-1, 1)
    return np.meshgrid(v, u, indexing="ij")


def phantom(H: int, W: int, F: int, rng: np.random.Generator, n_ellipses: int = 10, cine: bool = True):
    """Sum of ellipses with intensity U(0.2,1) times a smooth linear phase; radii breathe sinusoidally over frames.

    Field of view is [-1,1)^2 (FOV = 2).  Centres U(-0.35,0.35)*FOV, semi-axes U(0.05,0.3)*FOV, angle U(0,pi).
    Cine: each ellipse's radii are scaled by (1 + 0.15*sin(2*pi*f/F + phi_e)).
    """
    V, U = _grid(H, W)
    fov = 2.0
    cx = rng.uniform(-0.35, 0.35, n_ellipses) * fov
    cy = rng.uniform(-0.35, 0.35, n_ellipses) * fov
    ax = rng.uniform(0.05, 0.3, n_ellipses) * fov
    ay = rng.uniform(0.05, 0.3, n_ellipses) * fov
    th = rng.uniform(0.0, np.pi, n_ellipses)
    inten = rng.uniform(0.2, 1.0, n_ellipses)
    phi = rng.uniform(0.0, 2 * np.pi, n_ellipses)
    alpha, beta = rng.uniform(-0.5, 0.5, 2)
    phase = np.exp(1j * np.pi * (alpha * U + beta * V))
    x = np.zeros((F, H, W), dtype=np.complex128)
    for f in range(F):
        img = np.zeros((H, W))
        for e in range(n_ellipses):
            s = 1.0 + (0.15 * np.sin(2 * np.pi * f / F + phi[e]) if cine else 0.0)
            du, dv = U - cx[e], V - cy[e]
            ct, st = np.cos(th[e]), np.sin(th[e])
            xr = (du * ct + dv * st) / (ax[e] * s)
            yr = (-du * st + dv * ct) / (ay[e] * s)
            img += inten[e] * ((xr * xr + yr * yr) <= 1.0)
        x[f] = img * phase
    return x


def sens_maps(H: int, W: int, C: int):
    """C complex Gaussian coil profiles (sigma = 0.4*FOV) centred on a ring of radius 0.5*FOV at angle 2*pi*c/C with a
    constant phase exp(2*pi*i*c/C), normalised so that sqrt(sum_c |S_c|^2) = 1 at every pixel."""
    V, U = _grid(H, W)
    fov = 2.0
    sigma = 0.4 * fov
    S = np.zeros((C, H, W), dtype=np.complex128)
    for c in range(C):
        a = 2 * np.pi * c / C
        cu, cv = 0.5 * fov * np.cos(a), 0.5 * fov * np.sin(a)
        g = np.exp(-((U - cu) ** 2 + (V - cv) ** 2) / (2 * sigma * sigma))
        S[c] = g * np.exp(1j * a)
    rss = np.sqrt(np.sum(np.abs(S) ** 2, axis=0))
    return S / rss


def line_mask(H: int, W: int, F: int, accel: float, n_centre: int, rng: np.random.Generator):
    """Per-frame variable-density ky-line mask, unshifted: the n_centre centre lines are rows {0..n/2-1} and
    {H-n/2..H-1}; round(H/accel) - n_centre further rows are drawn without replacement with density proportional to
    (1 - |ky|/(H/2))^2, redrawn for every frame.  Rows are broadcast over kx."""
    half = n_centre // 2
    ky = np.where(np.arange(H) < H // 2, np.arange(H), np.arange(H) - H).astype(np.float64)
    centre = np.zeros(H, dtype=bool)
    centre[:half] = True
    centre[H - half:] = True
    n_lines = int(round(H / accel))
    n_extra = max(0, min(n_lines - n_centre, H - n_centre))
    M = np.zeros((F, H, W), dtype=np.uint8)
    cand = np.nonzero(~centre)[0]
    w = (1.0 - np.abs(ky[cand]) / (H / 2)) ** 2
    w = np.maximum(w, 1e-12)
    p = w / w.sum()
    for f in range(F):
        rows = centre.copy()
        if n_extra > 0:
            pick = rng.choice(cand, size=n_extra, replace=False, p=p)
            rows[pick] = True
        M[f, rows, :] = 1
    return M


def fft2_unscaled(a):
    return np.fft.fft2(a, axes=(-2, -1))


@dataclasses.dataclass
class Slice:
    """One reconstruction problem (one cine slice)."""
    y: np.ndarray          # [F][C][H][W] complex64 k-space
    S: np.ndarray          # [C][H][W] complex64
    M: np.ndarray          # [F][H][W] uint8
    x_true: np.ndarray     # [F][H][W] complex64 ground-truth image
    H: int
    W: int
    C: int
    F: int

    @property
    def dims(self):
        return (self.F, self.C, self.H, self.W)


def make_fully_sampled(H=256, W=256, C=8, F=1, seed=BASE_SEED, slice_id=0) -> Slice:
    """SimpleMRIRecon input: y = FFT2_unscaled(S*x) + n, fully sampled (mask of ones)."""
    rng = slice_rng(seed, slice_id)
    x = phantom(H, W, F, rng, cine=F > 1)
    S = sens_maps(H, W, C)
    k = fft2_unscaled(S[None, :, :, :] * x[:, None, :, :])
    sig = 1e-3 * np.abs(k).max()
    n = rng.normal(0.0, sig / np.sqrt(2), k.shape) + 1j * rng.normal(0.0, sig / np.sqrt(2), k.shape)
    y = (k + n).astype(np.complex64)
    M = np.ones((F, H, W), dtype=np.uint8)
    return Slice(y, S.astype(np.complex64), M, x.astype(np.complex64), H, W, C, F)


def make_cine(H=256, W=256, C=8, F=20, accel=6.0, n_centre=16, seed=BASE_SEED, slice_id=0, noise=1e-3) -> Slice:
    """Undersampled cine slice: y = M * (F_u(S*x) + n) with the unitary 2-D DFT F_u = FFT2/sqrt(H*W)."""
    rng = slice_rng(seed, slice_id)
    x = phantom(H, W, F, rng, cine=True)
    S = sens_maps(H, W, C)
    M = line_mask(H, W, F, accel, n_centre, rng)
    k = fft2_unscaled(S[None, :, :, :] * x[:, None, :, :]) / np.sqrt(H * W)
    sig = noise * np.abs(k).max()
    n = rng.normal(0.0, sig / np.sqrt(2), k.shape) + 1j * rng.normal(0.0, sig / np.sqrt(2), k.shape)
    y = (M[:, None, :, :] * (k + n)).astype(np.complex64)
    return Slice(y, S.astype(np.complex64), M, x.astype(np.complex64), H, W, C, F)


def batch_sizes(n_slices=64, seed=BASE_SEED, sizes=range(128, 321, 32), coils=(4, 32), frames=(8, 30)):
    """Uneven batch: per slice H=W drawn from `sizes`, C uniform in [4,32], F uniform in [8,30]."""
    rng = np.random.Generator(np.random.PCG64(splitmix64(seed ^ 0xBA7C4)))
    sizes = list(sizes)
    out = []
    for _ in range(n_slices):
        n = int(rng.choice(sizes))
        out.append((n, n, int(rng.integers(coils[0], coils[1] + 1)), int(rng.integers(frames[0], frames[1] + 1))))
    return out


def make_image_parts(H, W, C, F, accel=6.0, n_centre=16, seed=BASE_SEED, slice_id=0):
    """Host-side parts of a cine slice (phantom x, maps S, mask M) for k-space synthesis on the device:
    y = M * (F_u(S x) + n), n at 1e-3 of the peak |F_u(S x)|."""
    rng = slice_rng(seed, slice_id)
    x = phantom(H, W, F, rng, cine=True).astype(np.complex64)
    S = sens_maps(H, W, C).astype(np.complex64)
    M = line_mask(H, W, F, accel, n_centre, rng)
    return x, S, M


def device_kspace(ctx, x, S, M, noise=1e-3, seed=BASE_SEED):
    """Builds y = M * (FFT2(S x)/sqrt(N) + n) on the device with the library's own operators; returns the device
    arrays (y, S, M)."""
    F, H, W = x.shape
    C = S.shape[0]
    dx, dS, dM = ctx.to_device(x), ctx.to_device(S), ctx.to_device(M)
    y = ctx.empty((F, C, H, W))
    ctx.cep(dx, dS, y, F, C, in_coils=False, conj=False)
    ctx.fft2d(y, y)
    ctx.scal(y, 1.0 / np.sqrt(H * W))
    peak = ctx.reduce(y, "maxabs")[0]
    ctx.add_noise_masked(y, dM, noise * peak, seed)
    del dx
    return y, dS, dM


def identity_mesh(H, W):
    """Original mesh of GroupwiseRegistration.cpp:150-164: int32 [2][W][H], element [d][c][r] = (c + 1, r + 1)."""
    mesh = np.empty((2, W, H), dtype=np.int32)
    mesh[0] = (np.arange(W, dtype=np.int32) + 1)[:, None]
    mesh[1] = (np.arange(H, dtype=np.int32) + 1)[None, :]
    return mesh


def motion_fields(H, W, F, amp=1.5, roi=0.6, seed=BASE_SEED, slice_id=0):
    """Smooth per-frame deformations for motion compensation: sample positions xn float32 [2][F][H][W] (1-based;
    [0] column, [1] row) displaced by at most `amp` pixels, the identity mesh, and a centred bound box (1-based,
    [col0, row0, col1, row1] as the interpolator reads it) covering `roi` of each side, with room for the margins
    (H / 40, W / 40) and the displacement inside the frame."""
    rng = slice_rng(seed, slice_id ^ 0x3C)
    r = np.arange(H, dtype=np.float64)[:, None]
    c = np.arange(W, dtype=np.float64)[None, :]
    xn = np.empty((2, F, H, W), dtype=np.float32)
    ph = rng.uniform(0, 2 * np.pi, size=(F, 4))
    for f in range(F):
        t = 2 * np.pi * f / F
        dx = amp * np.sin(2 * np.pi * r / H + ph[f, 0]) * np.cos(2 * np.pi * c / W + ph[f, 1]) * np.cos(t)
        dy = amp * np.cos(2 * np.pi * r / H + ph[f, 2]) * np.sin(2 * np.pi * c / W + ph[f, 3]) * np.sin(t + 0.5)
        xn[0, f] = (c + 1 + dx).astype(np.float32)
        xn[1, f] = (r + 1 + dy).astype(np.float32)
    hr, hc = int(roi * H / 2), int(roi * W / 2)
    bb = np.array([W // 2 - hc + 1, H // 2 - hr + 1, W // 2 + hc, H // 2 + hr], dtype=np.uint32)
    return xn, identity_mesh(H, W), bb


def shifted_blobs(N, F, shifts, seed=3, n=14, sig=(3.5, 6.0)):
    """Registration input: F frames of one smooth texture (n Gaussian blobs of width sig about the centre) with
    frame f translated by shifts[f] = (columns, rows), complex64 [F][N][N] with a constant phase."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:N, 0:N].astype(np.float64)
    c = N / 2
    spec = [(c + rng.uniform(-N / 5, N / 5), c + rng.uniform(-N / 5, N / 5), rng.uniform(*sig), rng.uniform(0.3, 1))
            for _ in range(n)]
    out = []
    for f in range(F):
        sx, sy = shifts[f]
        img = np.zeros((N, N))
        for cy, cx, s, a in spec:
            img += a * np.exp(-((yy - cy - sy) ** 2 + (xx - cx - sx) ** 2) / (2 * s * s))
        out.append(img * np.exp(1j * 0.3))
    return np.array(out).astype(np.complex64)
