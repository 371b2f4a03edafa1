"""GPU parity of the split-precision NESTA-TV form (BF16-MFMA DFTs with a 3-term hi/lo split, FP32 accumulation,
FP64 scalars) against the FP64 oracle and the FP32 fused form, through the C ABI."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu
TOL = 1e-3   # rel-L2 bar of every form against the FP64 oracle


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def upload(ctx, sl):
    return (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))


@pytest.mark.parametrize("H,W,C,F,accel", [(64, 64, 4, 8, 4.0), (128, 64, 3, 5, 6.0), (192, 192, 4, 6, 8.0)])
def test_split_fixed_schedule_vs_fp64(ctx, oracle, H, W, C, F, accel):
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=accel, slice_id=31)
    prm = dict(maxIntIter=4, maxIter=12, stoptest=0)
    print("split case", H, W, C, F, accel, flush=True)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(ctx, sl)
    rep = ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    got = x.get()
    err = oracle.rel_l2(got, x64)
    ftr = np.asarray(rep["trace_f"]), np.asarray(r64["trace_f"])
    ferr = float(np.max(np.abs(ftr[0] - ftr[1]) / np.abs(ftr[1])))
    print(f"split {H}x{W} C{C} F{F}: rel-L2 vs FP64 {err:.2e}, max per-iteration |f - f64|/f64 {ferr:.2e}", flush=True)
    assert rep["iters"] == [12] * 4
    assert abs(rep["normU"] - r64["normU"]) < 1e-5 * r64["normU"]
    assert err < TOL
    assert ferr < TOL


def test_split_close_to_fp32_form(ctx, oracle):
    sl = synth.make_cine(H=128, W=128, C=6, F=10, accel=6.0, slice_id=32)
    prm = capi.nesta_params(maxIntIter=5, maxIter=10, stoptest=0)
    y, S, M, x = upload(ctx, sl)
    ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_FUSED)
    x32 = x.get()
    ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_SPLIT)
    xs = x.get()
    d = oracle.rel_l2(xs, x32)
    print(f"split vs FP32 fused form: {d:.2e}", flush=True)
    assert d < TOL


def test_split_stoptest_vs_fp64(ctx, oracle):
    sl = synth.make_cine(H=64, W=64, C=4, F=12, accel=4.0, slice_id=33)
    prm = dict(maxIntIter=4, maxIter=100, stoptest=1, miniter=7)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(ctx, sl)
    rep = ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    err = oracle.rel_l2(x.get(), x64)
    print(f"split stop test: iterations {rep['iters']} (FP64 oracle {r64['iters']}), rel-L2 {err:.2e}", flush=True)
    assert err < TOL


def test_split_batch_equals_solo(ctx, oracle):
    sls = [synth.make_cine(H=64, W=64, C=c, F=f, accel=4.0, slice_id=40 + i) for i, (c, f) in enumerate([(3, 6), (5, 9)])]
    prm = capi.nesta_params(maxIntIter=3, maxIter=8, stoptest=0)
    dev = [upload(ctx, s) for s in sls]
    bt = ctx.batch(dev, prm, form=capi.OCL_FORM_SPLIT)
    bt.run()
    got = [d[3].get() for d in dev]
    bt.close()
    for d, g in zip(dev, got):
        y, S, M, x = d
        ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_SPLIT)
        assert oracle.rel_l2(g, x.get()) < 1e-6


def test_split_rejects_unsupported_size(ctx):
    sl = synth.make_cine(H=160, W=160, C=2, F=3, accel=4.0, slice_id=34)
    y, S, M, x = upload(ctx, sl)
    with pytest.raises(capi.OclError) as e:
        ctx.nesta(y, S, M, x, capi.nesta_params(maxIntIter=1, maxIter=2, stoptest=0), form=capi.OCL_FORM_SPLIT)
    assert e.value.code == -3
