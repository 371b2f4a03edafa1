"""GPU checks of ocl_nesta_run_noalloc, the allocation-free variant of ocl_nesta_run: its output must equal
ocl_nesta_run's on the same inputs, call after call (the shared scratch array keeps the previous call's contents, so
the second call is smaller than the first), and a reconstruction that does not fit the scratch array is refused."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctxs():
    a, b = capi.Context(0), capi.Context(0)    # a: allocation-free calls, b: ocl_nesta_run (own normest, own cache)
    yield a, b
    a.close()
    b.close()


def _run_pair(ctxs, sl, prm, form):
    a, b = ctxs
    outs, reps = [], []
    for c, fn in ((a, a.nesta_noalloc), (b, b.nesta)):
        y, S, M = c.to_device(sl.y), c.to_device(sl.S), c.to_device(sl.M)
        x = c.empty((sl.F, sl.H, sl.W))
        reps.append(fn(y, S, M, x, capi.nesta_params(**prm), form=form))
        outs.append(x.get())
    return outs, reps


@pytest.mark.parametrize("form", [capi.OCL_FORM_FUSED, capi.OCL_FORM_SPLIT])
def test_noalloc_equals_nesta_run_two_calls(ctxs, oracle, form):
    prm = dict(maxIntIter=3, maxIter=8, stoptest=0)
    cases = [dict(H=128, W=128, C=8, F=20, accel=6.0, slice_id=41),     # first call: the larger one
             dict(H=64, W=64, C=4, F=8, accel=4.0, slice_id=42)]        # second call: smaller, on a used scratch
    for k, cs in enumerate(cases):
        print("noalloc form", form, "call", k, cs, flush=True)
        sl = synth.make_cine(**cs)
        (xa, xb), (ra, rb) = _run_pair(ctxs, sl, prm, form)
        assert ra["iters"] == rb["iters"] == [8] * 3
        assert ra["normU"] == rb["normU"]
        np.testing.assert_array_equal(xa, xb)
        np.testing.assert_array_equal(ra["trace_f"], rb["trace_f"])
        x64, _ = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
        assert oracle.rel_l2(xa, x64) < (1e-4 if form == capi.OCL_FORM_FUSED else 1e-3)


def test_noalloc_stop_test_schedule(ctxs, oracle):
    """The reference's stop test (NestaUp.cpp:799-824): iteration counts equal ocl_nesta_run's and the oracle's."""
    prm = dict(maxIntIter=4, maxIter=100, stoptest=1, miniter=7)
    sl = synth.make_cine(H=64, W=64, C=4, F=12, accel=4.0, slice_id=43)
    print("noalloc stop test 12x4x64x64", flush=True)
    (xa, xb), (ra, rb) = _run_pair(ctxs, sl, prm, capi.OCL_FORM_FUSED)
    _, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**prm), double=True)
    assert ra["iters"] == rb["iters"] == r64["iters"]
    np.testing.assert_array_equal(xa, xb)


def test_noalloc_refuses_what_does_not_fit(ctxs):
    a, _ = ctxs
    F, C, H, W = 160, 2, 256, 256          # 4 image-sized state arrays alone take 320 MiB
    y, S = a.empty((F, C, H, W)), a.empty((C, H, W))
    M = a.to_device(np.ones((F, H, W), np.uint8))
    x = a.empty((F, H, W))
    with pytest.raises(capi.OclError) as e:
        a.nesta_noalloc(y, S, M, x, capi.nesta_params(maxIntIter=1, maxIter=1, stoptest=0))
    assert e.value.code == capi.OCL_E_NOMEM
    assert "MiB scratch array" in str(e.value)
    with pytest.raises(capi.OclError) as e:
        a.nesta_noalloc(y, S, M, x, capi.nesta_params(), form=capi.OCL_FORM_PLAIN)
    assert e.value.code == capi.OCL_E_UNSUPPORTED
