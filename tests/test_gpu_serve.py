"""Served form: continuous batching with admission by HBM budget (ocl_server_*), through the C ABI."""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def test_served_results_equal_solo_and_budget_is_respected(ctx):
    shapes = [(128, 128, 4, 8, 10), (96, 160, 3, 6, 14), (64, 64, 2, 12, 8), (160, 128, 5, 6, 12),
              (128, 128, 4, 8, 6), (12, 10, 3, 4, 9), (256, 256, 8, 10, 10), (64, 96, 2, 5, 11)]
    slices = [synth.make_cine(H=h, W=w, C=c, F=f, accel=3.0, n_centre=4, slice_id=60 + i)
              for i, (h, w, c, f, _) in enumerate(shapes)]
    prms = [capi.nesta_params(maxIntIter=3, maxIter=it, stoptest=1 if i % 2 else 0, miniter=5)
            for i, (_, _, _, _, it) in enumerate(shapes)]
    dev = [(ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))
           for sl in slices]
    # a budget that holds the largest request but not all of them: later requests must wait for memory
    budget = 150 << 20
    srv = capi.Server(0, budget_bytes=budget, chunk=4)
    ids = [srv.submit(*dev[i], prms[i], want_report=True) for i in range(len(shapes))]
    for rid in ids:
        assert srv.wait(rid, 300.0) == capi.SERVE_DONE
    st = srv.stats()
    print(st, flush=True)
    assert st["completed"] == len(shapes) and st["failed"] == 0
    assert 0 < st["peak_used_bytes"] <= budget
    admits = [srv.times(r)[1] for r in ids]
    submits = [srv.times(r)[0] for r in ids]
    assert max(a - s for a, s in zip(admits, submits)) > 0      # somebody queued for memory
    for i, sl in enumerate(slices):
        y, S, M, x = (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))
        rs = ctx.nesta(y, S, M, x, prms[i], form=capi.OCL_FORM_FUSED)
        rb = srv.report(ids[i])
        print(i, shapes[i], rb["iters"], rs["iters"], flush=True)
        assert rb["iters"] == rs["iters"]
        assert np.array_equal(dev[i][3].get(), x.get())
    srv.close()


def test_reused_request_buffers_give_solo_results(ctx):
    """Requests of one shape reuse the work buffers of the ones before them (which still hold their data), and a
    request of another shape in between makes the server free cached buffers to stay inside its budget.  Every
    result must equal its solo run bit for bit."""
    a = synth.make_cine(H=128, W=128, C=4, F=8, accel=3.0, n_centre=4, slice_id=80)
    b = synth.make_cine(H=96, W=160, C=3, F=6, accel=3.0, n_centre=4, slice_id=81)
    order = [a, a, b, a, b, b, a]
    prm = capi.nesta_params(maxIntIter=3, maxIter=8, stoptest=0)
    ins = {id(sl): (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M)) for sl in (a, b)}
    outs = [ctx.empty((sl.F, sl.H, sl.W)) for sl in order]
    budget = 10 << 20          # holds one request (about 6.6 MiB and 4.4 MiB) at a time, not two
    srv = capi.Server(0, budget_bytes=budget, chunk=4)
    try:
        ids = [srv.submit(*ins[id(sl)], outs[i], prm) for i, sl in enumerate(order)]
        for rid in ids:
            assert srv.wait(rid, 300.0) == capi.SERVE_DONE
        st = srv.stats()
    finally:
        srv.close()
    print(st, flush=True)
    assert st["completed"] == len(order) and st["failed"] == 0
    assert st["cache_hits"] > 0 and st["peak_used_bytes"] <= budget
    solo = {}
    for sl in (a, b):
        x = ctx.empty((sl.F, sl.H, sl.W))
        ctx.nesta(*ins[id(sl)], x, prm, form=capi.OCL_FORM_FUSED)
        solo[id(sl)] = x.get()
    for i, sl in enumerate(order):
        assert np.array_equal(outs[i].get(), solo[id(sl)]), i


def test_server_rejects_request_larger_than_budget(ctx):
    sl = synth.make_cine(H=128, W=128, C=4, F=8, accel=3.0, slice_id=70)
    dev = (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))
    srv = capi.Server(0, budget_bytes=1 << 20, chunk=4)
    rid = srv.submit(*dev, capi.nesta_params(maxIntIter=2, maxIter=4, stoptest=0))
    with pytest.raises(capi.OclError) as e:
        srv.wait(rid, 60.0)
    assert e.value.code == -4
    srv.close()


def test_device_noise_helper_statistics(ctx):
    F, C, H, W = 2, 3, 64, 64
    M = np.ones((F, H, W), np.uint8)
    M[:, ::2] = 0
    y = ctx.zeros((F, C, H, W))
    ctx.add_noise_masked(y, ctx.to_device(M), 0.5, 1234)
    v = y.get()
    on = v[:, :, 1::2]
    assert np.all(v[:, :, ::2] == 0)
    assert abs(np.mean(np.abs(on) ** 2) - 0.25) < 0.01       # E|sigma n|^2 = sigma^2
    assert abs(np.mean(on.real)) < 0.01 and abs(np.mean(on.imag)) < 0.01


def test_server_refuses_requests_after_a_device_error(ctx):
    """After a device error the worker fails every running and queued request and stops; a later submission is
    refused at once (it would otherwise wait forever).  The error is injected from the host, so no kernel fails."""
    sl = synth.make_cine(H=64, W=64, C=2, F=4, accel=3.0, slice_id=71)
    dev = (ctx.to_device(sl.y), ctx.to_device(sl.S), ctx.to_device(sl.M), ctx.empty((sl.F, sl.H, sl.W)))
    srv = capi.Server(0, budget_bytes=1 << 30, chunk=2)
    try:
        rid = srv.submit(*dev, capi.nesta_params(maxIntIter=5, maxIter=200, stoptest=0))
        srv.inject_failure("injected for the test")
        with pytest.raises(capi.OclError) as e:
            srv.wait(rid, 60.0)
        assert e.value.code == -2 and "injected" in str(e.value)
        with pytest.raises(capi.OclError) as e:
            srv.submit(*dev, capi.nesta_params(maxIntIter=1, maxIter=2, stoptest=0))
        assert e.value.code == -2 and "stopped after a device error" in str(e.value)
        st = srv.stats()
        assert st["failed"] == 1 and st["completed"] == 0
    finally:
        srv.close()
