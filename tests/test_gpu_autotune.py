"""Tile-shape autotuner of the FP32 fused passes: every variant it can choose reconstructs within the parity bar,
and batches on an autotuning context use the tuned shape."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def run(ctx, sl, prm):
    y, S, M = ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M)
    x = ctx.empty((sl.F, sl.H, sl.W))
    rep = ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_FUSED)
    return x.get(), rep


# (ca_fast, tile_cols, b_fast, rows_per_wg) on a 160 x 96 slice: H = 16 * 10 and W = 16 * 6 allow every variant
# (ca_fast 3: the Gram form, with pass B in four-step or Stockham form for its setup)
VARIANTS = [(1, 16, 1, 16), (2, 16, 1, 16), (0, 4, 0, 4), (0, 8, 1, 16), (1, 16, 0, 8), (2, 16, 0, 16),
            (0, 16, 0, 16), (3, 16, 1, 16), (3, 16, 0, 8)]


@pytest.mark.parametrize("v", VARIANTS)
def test_every_tile_variant_matches_oracle(ctx, oracle, v):
    sl = synth.make_cine(H=160, W=96, C=3, F=5, accel=5.0, slice_id=51)
    prm = dict(maxIntIter=3, maxIter=8, stoptest=0)
    ref, _ = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    print("tile variant", v, flush=True)
    ctx.set_tiles(160, 96, *v)
    got, rep = run(ctx, sl, capi.nesta_params(**prm))
    err = oracle.rel_l2(got, ref)
    print(f"variant {v}: rel-L2 vs FP64 {err:.2e}", flush=True)
    assert rep["iters"] == [8] * 3
    assert err < 1e-4


def test_autotune_picks_a_valid_shape_and_is_used(ctx, oracle):
    c2 = capi.Context(0)
    try:
        t = c2.autotune_shape(192, 128)
        print("tuned 192x128:", t, flush=True)
        # the Gram form (ca_fast 3) runs no row pass per iteration: its ms_B is 0
        assert t["ms_CA"] > 0 and (t["ms_B"] > 0) == (t["ca_fast"] != 3)
        assert (t["ca_fast"], t["tile_cols"]) in {(1, 16), (2, 16), (0, 4), (0, 8), (0, 16), (3, 16)}
        assert (t["b_fast"], t["rows_per_wg"]) in {(1, 16), (0, 4), (0, 8), (0, 16)}
        assert c2.autotune_shape(192, 128) == t          # cached
        c2.set_autotune(True)
        sl = synth.make_cine(H=224, W=224, C=4, F=6, accel=6.0, slice_id=52)
        prm = dict(maxIntIter=3, maxIter=6, stoptest=0)
        ref, _ = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
        got, _ = run(c2, sl, capi.nesta_params(**prm))
        assert oracle.rel_l2(got, ref) < 1e-4
        t2 = c2.autotune_shape(224, 224)                  # tuned during the batch setup above
        print("tuned 224x224:", t2, flush=True)
        assert t2["ms_CA"] > 0
    finally:
        c2.close()


def test_tuned_shapes_are_shared_by_contexts_on_a_device():
    """A second context adopts the shape the first one tuned (the same kernels, so bit-identical results), while a
    shape pinned on one context stays on that context."""
    a, b = capi.Context(0), capi.Context(0)
    try:
        ta = a.autotune_shape(160, 160)
        assert b.autotune_shape(160, 160) == ta
        b.set_tiles(128, 128, 0, 8, 0, 8)
        tb = b.autotune_shape(128, 128)
        assert (tb["ca_fast"], tb["tile_cols"], tb["b_fast"], tb["rows_per_wg"]) == (0, 8, 0, 8)
        assert tb["ms_CA"] == 0.0                          # pinned, not timed
        assert a.autotune_shape(128, 128)["ms_CA"] > 0     # a tuned entry, not b's pinned one
    finally:
        a.close()
        b.close()


def test_set_tiles_rejects_bad_shapes(ctx):
    with pytest.raises(capi.OclError):
        ctx.set_tiles(100, 96, 1, 16, 1, 16)      # 100 is not 16 P
    with pytest.raises(capi.OclError):
        ctx.set_tiles(160, 96, 0, 3, 1, 16)       # not a power of two


@pytest.mark.parametrize("H,W", [(320, 96), (288, 64)])
def test_many_sampled_rows_take_the_direct_column_pass(oracle, H, W):
    """More than 64 sampled rows per frame (acceleration 2) exceed the prefetching pass's staging slabs, so the
    four-step column pass gathers the sampled rows straight from global memory (P = 20: two DFT-16 rounds per thread);
    the result stays within the parity bar of the FP64 oracle, and both four-step settings agree bit for bit."""
    sl = synth.make_cine(H=H, W=W, C=2, F=3, accel=2.0, n_centre=8, slice_id=57)
    assert sl.M.any(axis=2).sum(axis=1).max() > 64
    prm = dict(maxIntIter=2, maxIter=6, stoptest=0)
    ref, _ = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    outs = []
    for ca_fast in (1, 2):
        c = capi.Context(0)
        try:
            c.set_tiles(H, W, ca_fast, 16, 1, 16)
            print("direct column pass", H, W, "ca_fast", ca_fast, flush=True)
            got, _ = run(c, sl, capi.nesta_params(**prm))
        finally:
            c.close()
        err = oracle.rel_l2(got, ref)
        print(f"rel-L2 vs FP64 {err:.2e}", flush=True)
        assert err < 1e-4
        outs.append(got)
    assert np.array_equal(outs[0], outs[1])
