"""GPU parity of groupwise B-spline registration (GroupwiseRegistration, src/processes/GroupwiseRegistration.cpp,
groupwiseRegistration/optimizer.cpp) through the C ABI, against the kernel-by-kernel CPU restatement
(oracle/gwr_oracle.py), and the handoff of its result to motion compensation.

Tolerances: the device works in FP32 with FP64 block sums, the oracle in FP64; over tens of gradient steps the cost
trajectory agrees to 1e-4 relative and the coefficients to 2e-3 relative L2.  Each case is chosen so that no step's
cost change is within rounding of zero (the adaptive step takes the other branch there)."""
import numpy as np
import pytest

from opencliper_amd import capi, synth
from oracle import gwr_oracle as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _run_gpu(ctx, x, prm, T0=None, registered=True):
    F, N, _ = x.shape
    reg = ctx.registration(F, N, N, prm)
    dx = ctx.to_device(x)
    T = ctx.empty((2, F, reg.Cp, reg.Cp), np.float32)
    if T0 is not None:
        T.set(np.asarray(T0, np.float32).reshape(2, F, reg.Cp, reg.Cp))
    xn = ctx.empty((2, F, N, N), np.float32)
    rg = ctx.empty((F, N, N), np.float32) if registered else None
    rep = reg.run(dx, T, xn, T_in=T0 is not None, registered=rg)
    out = {"rep": rep, "T": T.get(), "xn": xn.get(), "reg": reg, "registered": rg.get() if rg is not None else None}
    return out


def _no_tie(want):
    """Every step's cost change is well away from zero relative to the cost itself."""
    return (np.abs(want["Hdif"]) / want["H"][1:]).min() > 1e-3


def _check_against_oracle(got, want, F, N, t_tol=2e-3):
    rep = got["rep"]
    assert rep["iters"] == want["iters"]
    n = want["iters"]
    assert np.allclose(rep["H"][: n + 1], want["H"][: n + 1], rtol=1e-4, atol=0), (rep["H"][:6], want["H"][:6])
    assert np.allclose(rep["Wn_trace"][:n], want["Wn"][:n], rtol=1e-6)
    T = got["T"].ravel().astype(np.float64)
    assert np.linalg.norm(T - want["T"]) <= t_tol * np.linalg.norm(want["T"])
    assert np.abs(got["xn"].ravel() - want["xn"]).max() < 2e-3
    if got["registered"] is not None:
        r = got["registered"].ravel().astype(np.float64)
        assert np.linalg.norm(r - want["transformed"]) <= 1e-4 * np.linalg.norm(want["transformed"])


# nmax stops each case before the step weight, grown x1.2 per step, overshoots: past that the trajectory is unstable
# and FP32 rounding is amplified into visible differences (e.g. after 17 steps of the first case)
@pytest.mark.parametrize("W,nmax,seed", [(0.05, 15, 0), (0.2, 12, 1)])
def test_registration_matches_oracle(ctx, W, nmax, seed):
    N, F = 64, 6
    print("gwr", N, F, W, nmax, seed, flush=True)
    rng = np.random.default_rng(seed)
    x = synth.shifted_blobs(N, F, rng.uniform(-0.8, 0.8, (F, 2)), seed=9 + seed, n=24)
    prm = dict(W=W, nmax=nmax, radius=16)
    got = _run_gpu(ctx, x, capi.gwr_params(**prm))
    want = G.Registration(N, F, **prm).run(x)
    assert _no_tie(want)
    _check_against_oracle(got, want, F, N)
    reg = got["reg"]
    assert reg.bound_box == want["bound_box"] and reg.Cp == want["Cp"]
    assert np.array_equal(reg.mesh().ravel(), want["x"])


def test_registration_reference_defaults_and_restart(ctx):
    """tests/MRIRecon.cpp:122-130 parameters (W = 1, lambda = {0, 0.005, 0, 0.5}) at 96^2, then a second launch that
    starts from the first one's coefficients (LaunchParameters::TData of the next motion iteration)."""
    N, F = 96, 5
    rng = np.random.default_rng(7)
    x = synth.shifted_blobs(N, F, rng.uniform(-1, 1, (F, 2)), seed=21, n=30)
    prm = dict(nmax=6, radius=24)
    got = _run_gpu(ctx, x, capi.gwr_params(**prm))
    O = G.Registration(N, F, **prm)
    want = O.run(x)
    assert _no_tie(want)
    _check_against_oracle(got, want, F, N)
    got2 = _run_gpu(ctx, x, capi.gwr_params(**prm), T0=got["T"])
    want2 = O.run(x, T=want["T"])
    assert _no_tie(want2)
    _check_against_oracle(got2, want2, F, N, t_tol=4e-3)


def test_registration_with_regularisation_in_the_cost(ctx):
    """lambda whose int-accumulated sum is 1: sumLambdaMetric (optimizer.cl:138-187) enters the cost."""
    N, F = 64, 4
    rng = np.random.default_rng(3)
    x = synth.shifted_blobs(N, F, rng.uniform(-0.8, 0.8, (F, 2)), seed=4, n=20)
    prm = dict(W=0.05, nmax=8, radius=16, lam=(1.5, 0.01, 0.02, 0.05))
    got = _run_gpu(ctx, x, capi.gwr_params(**prm))
    assert got["reg"].use_lambda
    want = G.Registration(N, F, **prm).run(x)
    assert _no_tie(want)
    _check_against_oracle(got, want, F, N)


def test_registration_at_cine_size_feeds_motion_compensation(ctx):
    """256^2 x 20 frames, radius 40 (the MRIRecon configuration): the cost falls, the stop rule ends the run, and
    MotionCompensation with the returned xn, mesh and bound box reproduces the registration's own resampled frames."""
    N, F = 256, 20
    rng = np.random.default_rng(11)
    x = synth.shifted_blobs(N, F, rng.uniform(-1.5, 1.5, (F, 2)), seed=5, n=40, sig=(6, 12))
    prm = capi.gwr_params(W=0.05, nmax=100)
    got = _run_gpu(ctx, x, prm)
    rep, reg = got["rep"], got["reg"]
    print("cine registration iters", rep["iters"], "H", rep["H"][0], "->", rep["H"][rep["iters"]], flush=True)
    assert reg.Cp == 23 and reg.support == 17 and reg.bound_box == [89, 89, 168, 168]
    assert rep["H"][rep["iters"]] < 0.2 * rep["H"][0]
    assert np.isfinite(got["T"]).all() and np.isfinite(got["xn"]).all()
    # the normalised frames through MotionCompensation at the estimated mesh = the registration's resampled frames
    I0 = (np.float32(500) / np.float32(rep["oldmax"]) * np.abs(x)).astype(np.float32)
    dI, out = ctx.to_device(I0.astype(np.complex64)), ctx.empty((F, N, N))
    dxn = ctx.to_device(got["xn"])
    ctx.motion_compensation(dI, out, reg.mc_args(dxn))
    mc = out.get()
    assert np.abs(mc.imag).max() == 0
    assert np.abs(mc.real - got["registered"]).max() <= 1e-4 * np.abs(I0).max()


def test_registration_error_paths(ctx):
    with pytest.raises(capi.OclError) as e:
        ctx.registration(4, 64, 80)
    assert e.value.code == capi.OCL_E_UNSUPPORTED and "square" in str(e.value)
    for kw in (dict(radius=0), dict(E=2), dict(Dp=(4, 5)), dict(radius=30)):
        with pytest.raises(capi.OclError) as e:
            ctx.registration(4, 64, 64, capi.gwr_params(**kw))
        assert e.value.code == capi.OCL_E_UNSUPPORTED, kw
    with pytest.raises(capi.OclError):
        ctx.registration(1, 64, 64, capi.gwr_params(radius=16))
    reg = ctx.registration(4, 64, 64, capi.gwr_params(radius=16))
    z = ctx.to_device(np.zeros((4, 64, 64), np.complex64))
    T = ctx.empty((2, 4, reg.Cp, reg.Cp), np.float32)
    xn = ctx.empty((2, 4, 64, 64), np.float32)
    with pytest.raises(capi.OclError) as e:
        reg.run(z, T, xn)
    assert e.value.code == capi.OCL_E_INVALID
