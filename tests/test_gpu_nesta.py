"""GPU parity of NESTA-TV (NestaUp::launch) in the plain and fused forms, through the C ABI."""
import json
import os

import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = sorted(json.load(open(os.path.join(GOLD, "golden.json"))).keys())


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def upload(ctx, y, S, M):
    F, C, H, W = y.shape
    return ctx.to_device(y.astype(np.complex64)), ctx.to_device(S.astype(np.complex64)), ctx.to_device(
        M.astype(np.uint8)), ctx.empty((F, H, W))


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
@pytest.mark.parametrize("name", NAMES)
def test_nesta_golden(ctx, oracle, name, form):
    info = json.load(open(os.path.join(GOLD, "golden.json")))[name]
    d = np.load(os.path.join(GOLD, f"{name}.npz"))
    print("case", name, "form", form, flush=True)
    y, S, M, x = upload(ctx, d["y"], d["S"], d["M"])
    rep = ctx.nesta(y, S, M, x, capi.nesta_params(**info["nesta"]), form=form)
    assert rep["iters"] == info["iters"]
    assert abs(rep["normU"] - info["normU"]) < 2e-6 * info["normU"]
    assert oracle.rel_l2(x.get(), d["x"]) < 1e-4


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
def test_nesta_medium_fixed_iterations(ctx, oracle, form):
    """F=20, C=8, 128x128, R=6, 5 stages x 10 iterations, stop test off."""
    sl = synth.make_cine(H=128, W=128, C=8, F=20, accel=6.0, slice_id=11)
    prm = dict(maxIntIter=5, maxIter=10, stoptest=0)
    x32, r32 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=False)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=form)
    got = x.get()
    print("form", form, "vs fp32", oracle.rel_l2(got, x32), "vs fp64", oracle.rel_l2(got, x64), flush=True)
    assert rep["iters"] == [10] * 5
    assert oracle.rel_l2(got, x32) < 1e-4
    assert oracle.rel_l2(got, x64) < 1e-4
    np.testing.assert_allclose(rep["trace_f"], r64["trace_f"], rtol=1e-4)


def test_nesta_fused_stoptest_matches_oracle(ctx, oracle):
    """Reference-faithful schedule (4 stages x up to 100 iterations, stop test on) on a small cine slice."""
    sl = synth.make_cine(H=64, W=64, C=4, F=12, accel=4.0, slice_id=12)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(), double=True)
    for form in (capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED):
        y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
        rep = ctx.nesta(y, S, M, x, capi.nesta_params(), form=form)
        print("form", form, rep["iters"], r64["iters"], oracle.rel_l2(x.get(), x64), flush=True)
        assert rep["iters"] == r64["iters"]
        assert oracle.rel_l2(x.get(), x64) < 1e-3


def test_fused_is_deterministic(ctx):
    sl = synth.make_cine(H=96, W=96, C=4, F=10, accel=4.0, slice_id=13)
    outs = []
    for _ in range(2):
        y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
        ctx.nesta(y, S, M, x, capi.nesta_params(maxIntIter=3, maxIter=8, stoptest=0), form=capi.OCL_FORM_FUSED)
        outs.append(x.get())
    assert np.array_equal(outs[0], outs[1])


def test_nesta_fast_path_wide_rows(ctx, oracle):
    """H = 160 (column four-step P=10), W = 320 (row four-step P=20) against the oracle."""
    sl = synth.make_cine(H=160, W=320, C=4, F=6, accel=4.0, slice_id=14)
    prm = dict(maxIntIter=2, maxIter=6, stoptest=0)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    for form in (capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED):
        y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
        ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=form)
        err = oracle.rel_l2(x.get(), x64)
        print("form", form, "vs fp64", err, flush=True)
        assert err < 1e-4


def _solo(ctx, sl, prm):
    y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_FUSED)
    return x.get(), rep


@pytest.mark.parametrize("stoptest", [0, 1])
def test_batch_uneven_equals_solo(ctx, oracle, stoptest):
    """Uneven slices advanced together give bit-identical results to their solo runs (no cross-slice coupling)."""
    shapes = [(128, 128, 4, 8), (96, 160, 3, 5), (64, 64, 2, 12), (12, 10, 3, 4), (160, 128, 5, 6)]
    prm = capi.nesta_params(maxIntIter=3, maxIter=12, stoptest=stoptest, miniter=5)
    slices = [synth.make_cine(H=h, W=w, C=c, F=f, accel=3.0, n_centre=4, slice_id=20 + i)
              for i, (h, w, c, f) in enumerate(shapes)]
    dev = [upload(ctx, sl.y, sl.S, sl.M) for sl in slices]
    b = ctx.batch(dev, prm)
    b.run()
    for i, sl in enumerate(slices):
        xs, rs = _solo(ctx, sl, prm)
        xb = dev[i][3].get()
        rb = b.report(i)
        print(i, shapes[i], rb["iters"], rs["iters"], flush=True)
        assert rb["iters"] == rs["iters"]
        assert np.array_equal(xb, xs)
    x64, _ = oracle.nesta(slices[1].y, slices[1].S, slices[1].M,
                          oracle.params(maxIntIter=3, maxIter=12, stoptest=stoptest, miniter=5), double=True)
    assert oracle.rel_l2(dev[1][3].get(), x64) < 1e-3
    b.close()


def test_batch_graph_replay_matches_eager(ctx):
    """A captured hipGraph of the iterations, replayed after fresh setups, reproduces the eager result each time."""
    sl = synth.make_cine(H=128, W=128, C=4, F=8, accel=4.0, slice_id=30)
    prm = capi.nesta_params(maxIntIter=2, maxIter=5, stoptest=0)
    dev = [upload(ctx, sl.y, sl.S, sl.M)]
    b = ctx.batch(dev, prm)
    b.setup()
    b.iterate(10)
    b.finalize()
    ctx.sync()
    ref = dev[0][3].get()
    b.setup()
    b.capture(10)
    for rep in range(3):
        b.setup()
        b.replay()
        b.finalize()
        ctx.sync()
        assert np.array_equal(dev[0][3].get(), ref), rep
    b.close()


@pytest.mark.parametrize("form", [capi.OCL_FORM_FUSED, capi.OCL_FORM_SPLIT])
def test_many_frames(ctx, oracle, form):
    """F = 40: the temporal-TV reduction covers the frames in two groups and the periodic wrap crosses them."""
    sl = synth.make_cine(H=64, W=64, C=2, F=40, accel=4.0, slice_id=13)
    prm = dict(maxIntIter=3, maxIter=6, stoptest=0)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    y, S, M, x = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(y, S, M, x, capi.nesta_params(**prm), form=form)
    err = oracle.rel_l2(x.get(), x64)
    print("F=40 form", form, "rel-L2 vs FP64", err, flush=True)
    assert rep["iters"] == [6] * 3
    np.testing.assert_allclose(rep["trace_f"], r64["trace_f"], rtol=1e-4)
    assert err < 1e-4 if form == capi.OCL_FORM_FUSED else err < 1e-3


@pytest.mark.parametrize("H,W,F", [(128, 160, 5), (96, 288, 4), (64, 320, 3)])
def test_pixel_mask_within_sampled_rows(ctx, oracle, H, W, F):
    """A sampling mask that is not constant along kx (OpenCLIPER's pixel masks; only ROWMASK masks are whole rows):
    inside the sampled rows every pixel is kept with probability 0.7.  The fused row pass reads a thread's 16 mask
    values (kx = k1 + P k2) as the bits of one precomputed word; W = 16 P with P = 10, 18 and 20 covers one and two
    roles per thread.  Both forms stay within 1e-4 of the FP64 oracle."""
    sl = synth.make_cine(H=H, W=W, C=3, F=F, accel=4.0, n_centre=8, slice_id=90 + P_of(W))
    rng = np.random.default_rng(91)
    M = (sl.M.astype(bool) & (rng.random(sl.M.shape) < 0.7)).astype(np.uint8)
    assert not all((M[f].min(axis=1) == M[f].max(axis=1)).all() for f in range(F))
    y = (sl.y * M[:, None]).astype(np.complex64)
    prm = dict(maxIntIter=2, maxIter=8, stoptest=0)
    x64, _ = oracle.nesta(y, sl.S, M, oracle.params(**prm), double=True)
    for form in (capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED):
        print("pixel mask", H, W, F, "form", form, flush=True)
        dy, dS, dM, x = upload(ctx, y, sl.S, M)
        ctx.nesta(dy, dS, dM, x, capi.nesta_params(**prm), form=form)
        err = oracle.rel_l2(x.get(), x64)
        print(f"rel-L2 vs FP64 {err:.2e}", flush=True)
        assert err < 1e-4


def P_of(W):
    return W // 16


def test_pixel_mask_split_precision(ctx, oracle):
    """The same kind of pixel mask through the split-precision form (BF16-MFMA DFTs): within 1e-3 of FP64."""
    H, W, F = 128, 128, 4
    sl = synth.make_cine(H=H, W=W, C=3, F=F, accel=4.0, n_centre=8, slice_id=95)
    rng = np.random.default_rng(96)
    M = (sl.M.astype(bool) & (rng.random(sl.M.shape) < 0.7)).astype(np.uint8)
    y = (sl.y * M[:, None]).astype(np.complex64)
    prm = dict(maxIntIter=2, maxIter=8, stoptest=0)
    x64, _ = oracle.nesta(y, sl.S, M, oracle.params(**prm), double=True)
    dy, dS, dM, x = upload(ctx, y, sl.S, M)
    ctx.nesta(dy, dS, dM, x, capi.nesta_params(**prm), form=capi.OCL_FORM_SPLIT)
    err = oracle.rel_l2(x.get(), x64)
    print(f"split form, pixel mask: rel-L2 vs FP64 {err:.2e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("H,W,stoptest", [(128, 128, 0), (160, 96, 1), (288, 64, 0), (320, 80, 1), (288, 72, 1)])
def test_gram_form_matches_row_and_column_passes(oracle, H, W, stoptest):
    """With a line mask the fused form runs the Gram iteration (one column pass, no row transform per iteration); the
    same slice with the column and row passes pinned (ca_fast 2) must agree with it and with FP64, and, with the stop
    test on, take the oracle's iteration counts (a stage ended by the stop test recomputes its last iteration).
    H = 288 and 320 give the column pass more DFT jobs than roles (two rounds); 288 x 72 also has a partial column tile."""
    sl = synth.make_cine(H=H, W=W, C=3, F=7, accel=6.0, n_centre=8, slice_id=120 + H)
    prm = dict(maxIntIter=3, maxIter=40 if stoptest else 12, stoptest=stoptest, miniter=4)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    outs = []
    for pinned in (False, True):
        c = capi.Context(0)
        try:
            if pinned:                        # W = 80 has no four-step row pass: Stockham rows there
                c.set_tiles(H, W, 2, 16, *((1, 16) if W % 32 == 0 else (0, 16)))
            y, S, M, x = upload(c, sl.y, sl.S, sl.M)
            rep = c.nesta(y, S, M, x, capi.nesta_params(**prm), form=capi.OCL_FORM_FUSED)
            got = x.get()
        finally:
            c.close()
        err = oracle.rel_l2(got, x64)
        print(f"{H}x{W} stoptest {stoptest} {'row+column passes' if pinned else 'Gram form'}: iterations "
              f"{rep['iters']} (oracle {r64['iters']}), rel-L2 vs FP64 {err:.2e}", flush=True)
        assert rep["iters"] == r64["iters"]
        assert err < 1e-4
        np.testing.assert_allclose(rep["trace_f"], r64["trace_f"], rtol=1e-4)
        outs.append(got)
    # both are FP32 and within 1e-4 of FP64; they round differently (no row transforms in the Gram form), which
    # NESTA amplifies over long stages (about 7e-5 apart after a 40-iteration stage at 160 x 96)
    assert oracle.rel_l2(outs[0], outs[1]) < 2e-4
