"""GPU parity of the individual Processes (TemporalTV, VectorNormalization, ComplexElementProd, ApplyMask,
XImageSum, ComplexAbs(Pow2), CopyDataGPU, FFT, SimpleMRIRecon, BLAS-1 reductions, myNormest) against the oracle,
called through the C ABI."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    print("device:", c.info(), flush=True)
    yield c
    c.close()


def rnd(shape, seed, scale=1.0):
    r = np.random.default_rng(seed)
    return ((r.normal(size=shape) + 1j * r.normal(size=shape)) * scale).astype(np.complex64)


@pytest.mark.parametrize("F,N", [(1, 17), (2, 64), (7, 1000), (20, 65536)])
def test_ttv_bit_exact(ctx, oracle, F, N):
    x = rnd((F, N), F * N)
    d = ctx.to_device(x)
    o = ctx.empty((F, N))
    for fwd in (True, False):
        ctx.ttv(d, o, fwd)
        assert np.array_equal(o.get(), oracle.ttv(x, fwd, np.complex64))


def test_vecnorm(ctx, oracle):
    x = rnd((3, 5000), 1)
    d = ctx.to_device(x)
    o = ctx.empty(x.shape)
    ctx.vecnorm(d, o, 0.8)
    np.testing.assert_allclose(o.get(), oracle.vecnorm(x, np.float32(0.8)), rtol=2e-7, atol=1e-7)


@pytest.mark.parametrize("in_coils,conj", [(0, 0), (1, 1), (1, 0)])
def test_cep(ctx, oracle, in_coils, conj):
    F, C, N = 3, 4, 777
    S = rnd((C, N), 2)
    x = rnd((F, C, N) if in_coils else (F, N), 3)
    o = ctx.empty((F, C, N))
    ctx.cep(ctx.to_device(x), ctx.to_device(S), o, F, C, in_coils, conj)
    np.testing.assert_allclose(o.get(), oracle.cep(x, S, F, C, in_coils, conj), rtol=1e-6, atol=1e-6)


def test_mask_and_ximage_sum_exact(ctx, oracle):
    F, C, H, W = 4, 3, 16, 24
    k = rnd((F, C, H, W), 4)
    M = (np.random.default_rng(5).random((F, H, W)) > 0.5).astype(np.uint8)
    d = ctx.to_device(k)
    ctx.apply_mask(d, ctx.to_device(M))
    km = d.get()
    assert np.array_equal(km, oracle.apply_mask(k, M))
    o = ctx.empty((F, H, W))
    ctx.ximage_sum(d, o)
    assert np.array_equal(o.get(), oracle.ximage_sum(km))


def test_abs_copy(ctx):
    x = rnd((1000,), 6)
    d = ctx.to_device(x)
    o = ctx.empty(x.shape)
    ctx.complex_abs(d, o)
    np.testing.assert_allclose(o.get().real, np.abs(x.astype(np.complex128)), rtol=2e-7)
    assert np.all(o.get().imag == 0)
    ctx.complex_abs(d, o, pow2=True)
    np.testing.assert_allclose(o.get().real, np.abs(x.astype(np.complex128)) ** 2, rtol=3e-7)
    ctx.copy(d, o)
    assert np.array_equal(o.get(), x)


def test_pinned_host_upload_round_trip():
    """ocl_host_malloc memory: a page-locked copy holds the same values, uploads from it land unchanged, and it stays
    valid after its context is closed (it is freed with its last view, by ocl_host_free without a context)."""
    c = capi.Context(0)
    try:
        rng = np.random.default_rng(5)
        a = (rng.standard_normal((3, 4, 64, 96)) + 1j * rng.standard_normal((3, 4, 64, 96))).astype(np.complex64)
        m = (rng.random((3, 64, 96)) < 0.3).astype(np.uint8)
        pa, pm = c.pinned(a), c.pinned(m)
        assert pa.shape == a.shape and pa.dtype == a.dtype and np.array_equal(pa, a)
        assert np.array_equal(pm, m)
        da, dm = c.empty(a.shape), c.empty(m.shape, np.uint8)
        da.set(pa)
        dm.set(pm)
        assert np.array_equal(da.get(), a) and np.array_equal(dm.get(), m)
        view = pa[1:]
    finally:
        c.close()
    assert np.array_equal(view, a[1:])
    del pa, pm, view


@pytest.mark.parametrize("H,W", [(8, 8), (12, 10), (128, 128), (160, 192), (224, 256), (256, 256), (288, 320),
                                 (320, 288), (7, 5), (9, 49), (1, 64), (512, 96)])
def test_fft2d_vs_numpy(ctx, H, W):
    x = rnd((3, H, W), H * 1000 + W)
    d = ctx.to_device(x)
    o = ctx.empty(x.shape)
    ctx.fft2d(d, o)
    ref = np.fft.fft2(x.astype(np.complex128))
    assert np.abs(o.get() - ref).max() <= 3e-6 * np.abs(ref).max() * max(1.0, np.log2(H * W) / 8)
    ctx.fft2d(d, d, inverse=True)          # in place
    refi = np.fft.ifft2(x.astype(np.complex128))
    assert np.abs(d.get() - refi).max() <= 3e-6 * np.abs(refi).max() * max(1.0, np.log2(H * W) / 8)


def test_fft_rejects_unsupported_length(ctx):
    d = ctx.empty((1, 11, 8))
    with pytest.raises(capi.OclError) as e:
        ctx.fft2d(d, d)
    assert e.value.code == -3


def test_reductions(ctx):
    x = rnd((100003,), 7)
    y = rnd((100003,), 8)
    dx, dy = ctx.to_device(x), ctx.to_device(y)
    X, Y = x.astype(np.complex128), y.astype(np.complex128)
    assert abs(ctx.reduce(dx, "nrm2")[0] - np.linalg.norm(X)) < 1e-9 * np.linalg.norm(X)
    re, im = ctx.reduce(dx, "dotc", dy)
    ref = np.vdot(X, Y)
    assert abs(complex(re, im) - ref) < 1e-9 * abs(ref) + 1e-6
    assert abs(ctx.reduce(dx, "sum")[0] - (X.real.sum() + X.imag.sum())) < 1e-6
    assert abs(ctx.reduce(dx, "asum")[0] - (np.abs(X.real).sum() + np.abs(X.imag).sum())) < 1e-6
    assert abs(ctx.reduce(dx, "maxabs")[0] - np.abs(X).max()) < 1e-6


@pytest.mark.parametrize("F,N", [(3, 64), (4, 64), (20, 65536), (25, 36864), (7, 16384)])
def test_normest_matches_oracle(ctx, oracle, F, N):
    g, gi = ctx.normest(F, N)
    e, ei = oracle.normest(F, N, double=False)
    assert gi == ei
    assert abs(g - e) < 2e-6 * e


def test_normest_carried_stream_matches_oracle(oracle):
    """A context that carries the drand48 stream across estimates, as the reference's process does
    (NestaUp.cpp:375): two estimates in a row equal the restatement's first and continued ones, and the position
    counts the draws."""
    c = capi.Context(0)
    try:
        c.set_normest_stream(True)
        F, N = 20, 65536
        g1, i1 = c.normest(F, N)
        assert c.normest_stream_position() == F * N
        g2, i2 = c.normest(F, N)
        assert c.normest_stream_position() == 2 * F * N
    finally:
        c.close()
    e1, j1 = oracle.normest(F, N, double=False)
    e2, j2 = oracle.normest(F, N, double=False, drand_start=None)
    print(f"carried normest: {g1:.7f} {g2:.7f} (oracle {e1:.7f} {e2:.7f})", flush=True)
    assert (i1, i2) == (j1, j2)
    assert abs(g1 - e1) < 2e-6 * e1 and abs(g2 - e2) < 2e-6 * e2


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
def test_repeated_nesta_carries_the_stream(oracle, form):
    """Two NestaUp::launch calls in one process: the second one's myNormest continues the drand48 stream where the
    first left it, so its normU (and result) follow the restatement run with the carried stream."""
    sl = synth.make_cine(H=64, W=64, C=4, F=8, accel=4.0, slice_id=77)
    prm = dict(maxIntIter=3, maxIter=10, stoptest=0)
    c = capi.Context(0)
    try:
        c.set_normest_stream(True)
        y, S, M = c.to_device(sl.y), c.to_device(sl.S), c.to_device(sl.M)
        x = c.empty((sl.F, sl.H, sl.W))
        r1 = c.nesta(y, S, M, x, capi.nesta_params(**prm), form=form)
        r2 = c.nesta(y, S, M, x, capi.nesta_params(**prm), form=form)
        got = x.get()
        pos = c.normest_stream_position()
    finally:
        c.close()
    x1, o1 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    x2, o2 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True, drand_start=None)
    print(f"form {form}: normU {r1['normU']:.7f} then {r2['normU']:.7f} (oracle {o1['normU']:.7f} "
          f"{o2['normU']:.7f}), rel-L2 vs FP64 {oracle.rel_l2(got, x2):.2e}", flush=True)
    assert pos == o2["drand48_position"] == 2 * sl.F * sl.H * sl.W
    assert abs(r1["normU"] - o1["normU"]) < 2e-6 * o1["normU"]
    assert abs(r2["normU"] - o2["normU"]) < 2e-6 * o2["normU"]
    assert r2["normU"] != r1["normU"]
    assert oracle.rel_l2(got, x2) < 1e-4


@pytest.mark.parametrize("H,W,C,F", [(256, 256, 8, 1), (160, 96, 3, 2), (128, 256, 4, 9)])
def test_simple_mri_recon(ctx, oracle, H, W, C, F):
    sl = synth.make_fully_sampled(H=H, W=W, C=C, F=F)
    y = ctx.to_device(sl.y)
    S = ctx.to_device(sl.S)
    x = ctx.empty((F, H, W))
    ctx.simple_mri_recon(y, S, x)
    ref64 = oracle.simple_recon(sl.y, sl.S, double=True)
    got = x.get()
    assert oracle.rel_l2(got, ref64) < 1e-6
    assert oracle.rel_l2(got, oracle.simple_recon(sl.y, sl.S, double=False)) < 1e-6
    # like the reference, the input is left holding conj(S_c) * IFFT2(y_c) (SimpleMRIRecon.cpp:59-101)
    post = np.fft.ifft2(sl.y.astype(np.complex128), axes=(-2, -1)) * np.conj(sl.S.astype(np.complex128))[None]
    assert oracle.rel_l2(y.get(), post) < 1e-6


@pytest.mark.parametrize("H,W,C,F,n", [(256, 256, 8, 1, 70), (128, 160, 3, 2, 5), (96, 80, 2, 1, 3)])
def test_simple_mri_recon_batch(ctx, oracle, H, W, C, F, n):
    """The batched entry point: every input's result and in-place state as ocl_simple_mri_recon gives them; 70
    inputs span two launch pairs of 64; 96 x 80 (W = 16 P with P odd) takes the per-input fallback."""
    sls = [synth.make_fully_sampled(H=H, W=W, C=C, F=F, slice_id=300 + i) for i in range(n)]
    print("simple recon batch", H, W, C, F, n, flush=True)
    items = [(ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.empty((F, H, W))) for sl in sls]
    ctx.simple_mri_recon_batch(items)
    for i in (0, n // 2, n - 1):
        sl = sls[i]
        y, S, x = items[i]
        assert oracle.rel_l2(x.get(), oracle.simple_recon(sl.y, sl.S, double=True)) < 1e-6
        post = np.fft.ifft2(sl.y.astype(np.complex128), axes=(-2, -1)) * np.conj(sl.S.astype(np.complex128))[None]
        assert oracle.rel_l2(y.get(), post) < 1e-6
        y1, S1, x1 = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.empty((F, H, W))
        ctx.simple_mri_recon(y1, S1, x1)
        assert oracle.rel_l2(x.get(), x1.get()) < 1e-6


@pytest.mark.parametrize("H,W,C,F", [(256, 256, 8, 1), (96, 160, 5, 3)])
def test_rsos_and_simple_recon_sos(ctx, oracle, H, W, C, F):
    sl = synth.make_fully_sampled(H=H, W=W, C=C, F=F, slice_id=23)
    k = ctx.to_device(sl.y)
    out = ctx.empty((F, H, W))
    ctx.rsos(k, out)
    assert oracle.rel_l2(out.get(), oracle.rsos(sl.y)) < 1e-6
    x = ctx.empty((F, H, W))
    ctx.simple_mri_recon_sos(k, x)
    got = x.get()
    assert np.all(got.imag == 0)
    assert oracle.rel_l2(got, oracle.simple_recon_sos(sl.y)) < 1e-6
