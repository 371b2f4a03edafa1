"""GPU parity of motion compensation (MotionCompensation / AdjointMotionCompensation, optimizer.cl:832-930) and of
NESTA-TV with argsMC (NestaUp.cpp:201-278,483-491) through the C ABI, against the FP64 oracle."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _cplx(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _fields(ctx, H, W, F, amp, slice_id):
    xn, mesh, bb = synth.motion_fields(H, W, F, amp=amp, slice_id=slice_id)
    dxn, dmesh = ctx.to_device(xn), ctx.to_device(mesh)
    return (xn, mesh, bb), (dxn, dmesh, ctx.mc_args(dxn, dmesh, bb))


@pytest.mark.parametrize("H,W,F", [(96, 96, 8), (80, 112, 5)])
def test_mc_forward_and_adjoint_match_oracle(ctx, oracle, H, W, F):
    print("mc", H, W, F, flush=True)
    rng = np.random.default_rng(H + W + F)
    x = _cplx(rng, F, H, W)
    (xn, mesh, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=2.5, slice_id=H)
    dx, out = ctx.to_device(x), ctx.empty((F, H, W))
    ctx.motion_compensation(dx, out, mc)
    want = oracle.mc(x, xn, mesh, bb)
    assert oracle.rel_l2(out.get(), want) < 1e-6
    ctx.motion_compensation(dx, out, mc, adjoint=True)
    want_t = oracle.mc(x, xn, mesh, bb, adjoint=True)
    assert oracle.rel_l2(out.get(), want_t) < 1e-6


def test_mc_identity_is_exact_and_adjoint_identity(ctx, oracle):
    H, W, F = 64, 64, 6
    rng = np.random.default_rng(41)
    x, y = _cplx(rng, F, H, W), _cplx(rng, F, H, W)
    xn = np.empty((2, F, H, W), np.float32)
    xn[0] = np.arange(1, W + 1, dtype=np.float32)[None, None, :]
    xn[1] = np.arange(1, H + 1, dtype=np.float32)[None, :, None]
    mesh, bb = synth.identity_mesh(H, W), np.array([12, 10, 50, 52], np.uint32)
    dxn, dmesh = ctx.to_device(xn), ctx.to_device(mesh)
    mc = ctx.mc_args(dxn, dmesh, bb)
    dx, out = ctx.to_device(x), ctx.empty((F, H, W))
    for adjoint in (False, True):
        ctx.motion_compensation(dx, out, mc, adjoint=adjoint)
        assert np.array_equal(out.get(), x)
    # <MC x, y> = <x, MC^T y> under a real deformation
    (_, _, _), (dxn2, dmesh2, mc2) = _fields(ctx, H, W, F, amp=3.0, slice_id=9)
    dy, mx, mty = ctx.to_device(y), ctx.empty((F, H, W)), ctx.empty((F, H, W))
    ctx.motion_compensation(dx, mx, mc2)
    ctx.motion_compensation(dy, mty, mc2, adjoint=True)
    lhs = np.vdot(mx.get().astype(np.complex128), y.astype(np.complex128))
    rhs = np.vdot(x.astype(np.complex128), mty.get().astype(np.complex128))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


def test_mc_errors(ctx):
    H, W, F = 32, 32, 2
    (_, _, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.0, slice_id=10)
    dx, out = ctx.zeros((F, H, W)), ctx.empty((F, H, W))
    with pytest.raises(capi.OclError, match="alias"):
        ctx.motion_compensation(dx, dx, mc)
    bad = ctx.mc_args(dxn, dmesh, [1, 1, 16, 16])
    with pytest.raises(capi.OclError, match="leaves the frame"):
        ctx.motion_compensation(dx, out, bad)
    nul = capi.McArgs(None, dmesh.ptr)
    for i, v in enumerate(bb):
        nul.bound_box[i] = int(v)
    with pytest.raises(capi.OclError, match="null"):
        ctx.motion_compensation(dx, out, nul)


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
@pytest.mark.parametrize("H,W,C,F", [(48, 48, 4, 6), (64, 80, 3, 8), (128, 128, 4, 5)])
def test_nesta_mc_matches_oracle(ctx, oracle, H, W, C, F, form):
    """NESTA with motion compensation from the motion-free reconstruction, as tests/MRIRecon.cpp:240-272 chains
    them: within 1e-3 of the FP64 oracle with the same stage lengths, in the plain and the fused form (128 x 128
    takes the four-step column and row passes)."""
    print("nesta_mc", H, W, C, F, "form", form, flush=True)
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=50 + F)
    prm = dict(maxIntIter=3, maxIter=15, stoptest=1, miniter=7)
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
    x = ctx.empty((F, H, W))
    ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_PLAIN)
    x_ref = x.get()
    (xn, mesh, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.2, slice_id=11 + F)
    rep = ctx.nesta_mc(y, S, M, x, mc, capi.nesta_params(**prm), form=form)
    x64, r64 = oracle.nesta_mc(sl.y, sl.S, sl.M, x_ref, xn, mesh, bb, oracle.params(**prm), double=True)
    err = oracle.rel_l2(x.get(), x64)
    print("iters", rep["iters"], r64["iters"], "normU", rep["normU"], r64["normU"], "rel", err, flush=True)
    assert err < 1e-3
    assert rep["iters"] == r64["iters"]
    assert abs(rep["normU"] - r64["normU"]) < 1e-5 * r64["normU"]
    assert oracle.rel_l2(x.get(), x_ref) > 1e-4      # the deformation changed the solution


def test_fused_mc_deterministic_and_large_motion(ctx, oracle):
    """Small deformations take the gather form of the adjoint (deterministic: two runs agree bit for bit); a
    deformation of up to 5 px exceeds the gather radius and takes the atomic scatter, still within 1e-3 of FP64."""
    H, W, C, F = 128, 128, 3, 6
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=80)
    prm = dict(maxIntIter=2, maxIter=12, stoptest=1, miniter=7)
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
    x0 = sl.x_true.astype(np.complex64)
    for amp, sid in ((1.3, 81), (5.0, 82)):
        print("amp", amp, flush=True)
        (xn, mesh, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=amp, slice_id=sid)
        outs = []
        for _ in range(2):
            x = ctx.to_device(x0)
            ctx.nesta_mc(y, S, M, x, mc, capi.nesta_params(**prm), form=capi.OCL_FORM_FUSED)
            outs.append(x.get())
        x64, _ = oracle.nesta_mc(sl.y, sl.S, sl.M, x0, xn, mesh, bb, oracle.params(**prm), double=True)
        assert oracle.rel_l2(outs[0], x64) < 1e-3
        if amp < 2:
            assert np.array_equal(outs[0], outs[1])


def test_nesta_mc_form_errors(ctx):
    """The split form has no motion compensation; the fused one needs the original (identity) mesh."""
    H, W, C, F = 32, 32, 2, 4
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=60)
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
    x = ctx.to_device(sl.x_true.astype(np.complex64))
    (xn, mesh, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.0, slice_id=12)
    prm = capi.nesta_params(maxIntIter=2, maxIter=4)
    with pytest.raises(capi.OclError, match="plain or the fused form"):
        ctx.nesta_mc(y, S, M, x, mc, prm, form=capi.OCL_FORM_SPLIT)
    swapped = ctx.to_device(np.ascontiguousarray(mesh[::-1]))
    with pytest.raises(capi.OclError, match="identity"):
        ctx.nesta_mc(y, S, M, x, ctx.mc_args(dxn, swapped, bb), prm, form=capi.OCL_FORM_FUSED)
    ctx.nesta_mc(y, S, M, x, ctx.mc_args(dxn, swapped, bb), prm, form=capi.OCL_FORM_PLAIN)   # any mesh in plain


def test_fused_mc_batch_uneven_equals_solo(ctx, oracle):
    """Two slices of different shapes with their own deformations in one fused launch table (graph replay of the
    whole schedule) give each slice's solo fused result.  The adjoint's scatter adds with float atomics, whose order
    varies, so the comparison allows rounding-level differences."""
    prm = dict(maxIntIter=2, maxIter=10, stoptest=0, miniter=7)
    shapes = [(64, 64, 3, 6, 70), (96, 80, 2, 5, 71)]
    solo, batch_in, mcs = [], [], []
    for H, W, C, F, sid in shapes:
        sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=sid)
        y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
        x0 = synth.make_image_parts(H, W, C, F, slice_id=sid)[0]
        (_, _, _), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.0, slice_id=sid)
        xs = ctx.to_device(x0)
        ctx.nesta_mc(y, S, M, xs, mc, capi.nesta_params(**prm), form=capi.OCL_FORM_FUSED)
        solo.append(xs.get())
        xb = ctx.to_device(x0)
        batch_in.append((y, S, M, xb, dxn, dmesh))
        mcs.append(mc)
    bt = ctx.batch([b[:4] for b in batch_in], capi.nesta_params(**prm), mcs=mcs)
    bt.setup()
    bt.capture(prm["maxIntIter"] * prm["maxIter"])
    bt.replay()
    bt.finalize()
    ctx.sync()
    for i in range(len(shapes)):
        assert oracle.rel_l2(batch_in[i][3].get(), solo[i]) < 1e-5, i
