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


@pytest.mark.parametrize("H,W,C,F", [(48, 48, 4, 6), (64, 80, 3, 8)])
def test_nesta_mc_matches_oracle(ctx, oracle, H, W, C, F):
    """NESTA with motion compensation from the motion-free reconstruction, as tests/MRIRecon.cpp:240-272 chains
    them: within 1e-3 of the FP64 oracle with the same stage lengths."""
    print("nesta_mc", H, W, C, F, flush=True)
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=50 + F)
    prm = dict(maxIntIter=3, maxIter=15, stoptest=1, miniter=7)
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
    x = ctx.empty((F, H, W))
    ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_PLAIN)
    x_ref = x.get()
    (xn, mesh, bb), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.2, slice_id=11 + F)
    rep = ctx.nesta_mc(y, S, M, x, mc, capi.nesta_params(**prm))
    x64, r64 = oracle.nesta_mc(sl.y, sl.S, sl.M, x_ref, xn, mesh, bb, oracle.params(**prm), double=True)
    err = oracle.rel_l2(x.get(), x64)
    print("iters", rep["iters"], r64["iters"], "normU", rep["normU"], r64["normU"], "rel", err, flush=True)
    assert err < 1e-3
    assert rep["iters"] == r64["iters"]
    assert abs(rep["normU"] - r64["normU"]) < 1e-5 * r64["normU"]
    assert oracle.rel_l2(x.get(), x_ref) > 1e-4      # the deformation changed the solution


def test_nesta_mc_rejects_fused_form(ctx):
    H, W, C, F = 32, 32, 2, 4
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=3.0, n_centre=4, slice_id=60)
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M.astype(np.uint8))
    x = ctx.zeros((F, H, W))
    (_, _, _), (dxn, dmesh, mc) = _fields(ctx, H, W, F, amp=1.0, slice_id=12)
    with pytest.raises(capi.OclError, match="plain form"):
        ctx.nesta_mc(y, S, M, x, mc, capi.nesta_params(maxIntIter=2, maxIter=4), form=capi.OCL_FORM_FUSED)
