"""GPU parity of the split-precision NESTA-TV form (BF16-MFMA DFTs with a 3-term hi/lo split, FP32 accumulation,
FP64 scalars) against the FP64 oracle and the FP32 fused form, through the C ABI.  Both split-precision iterations are
covered: the Gram pass (column GEMMs only, the default for line masks) and the row and column passes (pinned with
ocl_ctx_set_split_gram(ctx, 0); also what pixel masks take)."""
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


@pytest.fixture(scope="module")
def ctx_rc():
    """A context whose split-precision batches keep the row and column passes."""
    c = capi.Context(0)
    c.set_split_gram(False)
    yield c
    c.close()


def upload(ctx, sl):
    return (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))


@pytest.mark.parametrize("gram", [True, False], ids=["gram", "rowcol"])
@pytest.mark.parametrize("H,W,C,F,accel", [(64, 64, 4, 8, 4.0), (128, 64, 3, 5, 6.0), (192, 192, 4, 6, 8.0)])
def test_split_fixed_schedule_vs_fp64(ctx, ctx_rc, oracle, gram, H, W, C, F, accel):
    c = ctx if gram else ctx_rc
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=accel, slice_id=31)
    prm = dict(maxIntIter=4, maxIter=12, stoptest=0)
    print("split case", "gram" if gram else "rowcol", H, W, C, F, accel, flush=True)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(c, sl)
    rep = c.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    got = x.get()
    err = oracle.rel_l2(got, x64)
    ftr = np.asarray(rep["trace_f"]), np.asarray(r64["trace_f"])
    ferr = float(np.max(np.abs(ftr[0] - ftr[1]) / np.abs(ftr[1])))
    print(f"split {H}x{W} C{C} F{F}: rel-L2 vs FP64 {err:.2e}, max per-iteration |f - f64|/f64 {ferr:.2e}", flush=True)
    assert rep["iters"] == [12] * 4
    assert abs(rep["normU"] - r64["normU"]) < 1e-5 * r64["normU"]
    assert err < TOL
    assert ferr < TOL


@pytest.mark.parametrize("H,W,C,F,accel", [(64, 64, 4, 8, 4.0), (192, 192, 5, 7, 8.0)])
def test_split_gram_and_rowcol_agree(ctx, ctx_rc, oracle, H, W, C, F, accel):
    """The two split-precision iterations round differently (so the Gram pass did run) but agree far inside the bar."""
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=accel, slice_id=35)
    prm = capi.nesta_params(maxIntIter=3, maxIter=10, stoptest=0)
    got = []
    for c in (ctx, ctx_rc):
        y, S, M, x = upload(c, sl)
        c.nesta(y, S, M, x, prm, form=capi.OCL_FORM_SPLIT)
        got.append(x.get())
    d = oracle.rel_l2(got[0], got[1])
    print(f"split Gram pass vs row and column passes {H}x{W}: {d:.2e}", flush=True)
    assert d > 0.0
    assert d < 1e-4


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


@pytest.mark.parametrize("gram", [True, False], ids=["gram", "rowcol"])
def test_split_stoptest_vs_fp64(ctx, ctx_rc, oracle, gram):
    c = ctx if gram else ctx_rc
    sl = synth.make_cine(H=64, W=64, C=4, F=12, accel=4.0, slice_id=33)
    prm = dict(maxIntIter=4, maxIter=100, stoptest=1, miniter=7)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(c, sl)
    rep = c.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    err = oracle.rel_l2(x.get(), x64)
    print(f"split stop test: iterations {rep['iters']} (FP64 oracle {r64['iters']}), rel-L2 {err:.2e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("gram", [True, False], ids=["gram", "rowcol"])
def test_split_batch_equals_solo(ctx, ctx_rc, oracle, gram):
    c = ctx if gram else ctx_rc
    sls = [synth.make_cine(H=64, W=64, C=cc, F=f, accel=4.0, slice_id=40 + i)
           for i, (cc, f) in enumerate([(3, 6), (5, 9)])]
    prm = capi.nesta_params(maxIntIter=3, maxIter=8, stoptest=0)
    dev = [upload(c, s) for s in sls]
    bt = c.batch(dev, prm, form=capi.OCL_FORM_SPLIT)
    bt.run()
    got = [d[3].get() for d in dev]
    bt.close()
    for d, g in zip(dev, got):
        y, S, M, x = d
        c.nesta(y, S, M, x, prm, form=capi.OCL_FORM_SPLIT)
        assert oracle.rel_l2(g, x.get()) < 1e-6


def test_split_mixed_forms_in_one_batch(ctx, oracle):
    """A split-precision batch holding a line-mask slice (Gram pass) and a pixel-mask slice (row and column passes)
    runs each form as a launch group of its own; each slice equals its solo run and stays within the bar."""
    a = synth.make_cine(H=64, W=64, C=3, F=6, accel=4.0, slice_id=46)
    b = synth.make_cine(H=64, W=64, C=4, F=5, accel=4.0, n_centre=8, slice_id=47)
    rng = np.random.default_rng(47)
    M = (b.M.astype(bool) & (rng.random(b.M.shape) < 0.7)).astype(np.uint8)
    b = synth.Slice((b.y * M[:, None]).astype(np.complex64), b.S, M, b.x_true, b.H, b.W, b.C, b.F)
    prm = dict(maxIntIter=3, maxIter=8, stoptest=0)
    dev = [upload(ctx, a), upload(ctx, b)]
    bt = ctx.batch(dev, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    bt.run()
    info = bt.info()
    got = [d[3].get() for d in dev]
    bt.close()
    print("mixed split batch:", info, flush=True)
    assert info["launch_groups"] == 2
    for sl, d, g in zip((a, b), dev, got):
        y, S, Mm, x = d
        ctx.nesta(y, S, Mm, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
        assert oracle.rel_l2(g, x.get()) < 1e-6
        x64, _ = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
        err = oracle.rel_l2(g, x64)
        print(f"slice {sl.F}x{sl.C}: rel-L2 vs FP64 {err:.2e}", flush=True)
        assert err < TOL


def test_split_rejects_unsupported_size(ctx):
    sl = synth.make_cine(H=160, W=160, C=2, F=3, accel=4.0, slice_id=34)
    y, S, M, x = upload(ctx, sl)
    with pytest.raises(capi.OclError) as e:
        ctx.nesta(y, S, M, x, capi.nesta_params(maxIntIter=1, maxIter=2, stoptest=0), form=capi.OCL_FORM_SPLIT)
    assert e.value.code == -3
