"""GPU parity at the benchmark workloads' exact sizes, through the C ABI (NestaUp::launch, NestaUp.cpp:442-910, with
the schedules of tests/MRIRecon.cpp:105-114):

  cine slice       20 frames x 8 coils x 256^2, acceleration 6, 5 stages x 20 iterations (stop test off) and the
                   reference's own schedule (100 iterations per stage, stop test on, miniter 7): plain, fused and
                   batched forms against the FP32 restatement and the FP64 oracle
  split precision  25 frames x 16 coils x 192^2, acceleration 8: BF16-MFMA DFT form against FP64
  uneven batch     the 64-slice table of 128..320 matrices, 4..32 coils, 8..30 frames: every slice equal to its solo
                   run, three of them against FP64
  launch groups    a mixed table whose 256^2 slices fill enough column-pass workgroups for a launch group of their own;
                   a one-size table with different coil counts, run longest first
  served trace     100 requests over the cine and uneven sizes, Poisson arrivals, an HBM budget that makes requests
                   wait: every served result equal to its solo run
"""
import numpy as np
import pytest

from opencliper_amd import capi, synth

pytestmark = pytest.mark.gpu

CINE = dict(H=256, W=256, C=8, F=20, accel=6.0)
FIXED = dict(maxIntIter=5, maxIter=20, stoptest=0)
REFERENCE = dict(maxIntIter=5, maxIter=100, stoptest=1, miniter=7)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def upload(ctx, y, S, M):
    F, C, H, W = y.shape
    return ctx.to_device(y), ctx.to_device(S), ctx.to_device(M), ctx.empty((F, H, W))


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def cine(oracle):
    """The headline slice (slice 0 of the benchmark batch) with its FP32 and FP64 references on the fixed schedule."""
    sl = synth.make_cine(**CINE, slice_id=0)
    print("oracle: cine 20x8x256x256, 5 x 20 iterations, FP32 and FP64", flush=True)
    x32, r32 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**FIXED), double=False)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**FIXED), double=True)
    return sl, x32, r32, x64, r64


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
def test_cine_fixed_schedule(ctx, oracle, cine, form):
    sl, x32, r32, x64, r64 = cine
    dev = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(*dev, capi.nesta_params(**FIXED), form=form)
    got = dev[3].get()
    e32, e64 = rel(got, x32), rel(got, x64)
    print(f"cine form {form}: rel-L2 vs FP32 restatement {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
    assert rep["iters"] == [20] * 5
    assert abs(rep["normU"] - r64["normU"]) < 2e-6 * r64["normU"]
    assert e32 < 1e-4 and e64 < 1e-4
    np.testing.assert_allclose(rep["trace_f"], r64["trace_f"], rtol=1e-4)


def test_cine_batch_of_16(ctx, oracle, cine):
    """The benchmark batch: 16 cine slices in one launch table (the homogeneous 256^2 kernels); slices 0 and 15
    against their FP32 and FP64 references."""
    sl0, x32, _, x64, _ = cine
    sl15 = synth.make_cine(**CINE, slice_id=15)
    print("oracle: slice 15, FP32 and FP64", flush=True)
    x32b, _ = oracle.nesta(sl15.y, sl15.S, sl15.M, oracle.params(**FIXED), double=False)
    x64b, _ = oracle.nesta(sl15.y, sl15.S, sl15.M, oracle.params(**FIXED), double=True)
    dev = [upload(ctx, sl0.y, sl0.S, sl0.M)]
    for i in range(1, 15):
        x, S, M = synth.make_image_parts(CINE["H"], CINE["W"], CINE["C"], CINE["F"], accel=CINE["accel"], slice_id=i)
        y, dS, dM = synth.device_kspace(ctx, x, S, M, seed=synth.BASE_SEED + i)
        dev.append((y, dS, dM, ctx.empty(x.shape)))
    dev.append(upload(ctx, sl15.y, sl15.S, sl15.M))
    b = ctx.batch(dev, capi.nesta_params(**FIXED))
    b.setup()
    info = b.info()
    b.capture(100)
    b.replay()
    b.finalize()
    ctx.sync()
    print("batch", info, flush=True)
    assert info["launch_groups"] == 0 and info["wg_CA"] == 16 * 20 * 16
    for i, (r32, r64) in ((0, (x32, x64)), (15, (x32b, x64b))):
        got = dev[i][3].get()
        e32, e64 = rel(got, r32), rel(got, r64)
        print(f"batched slice {i}: vs FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
        assert b.report(i)["iters"] == [20] * 5
        assert e32 < 1e-4 and e64 < 1e-4
    b.close()


@pytest.fixture(scope="module")
def cine_reference(oracle, cine):
    sl = cine[0]
    print("oracle: cine, stop-test schedule, FP32 and FP64", flush=True)
    return (oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**REFERENCE), double=False),
            oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**REFERENCE), double=True))


@pytest.mark.parametrize("form", [capi.OCL_FORM_PLAIN, capi.OCL_FORM_FUSED])
def test_cine_reference_schedule(ctx, cine, cine_reference, form):
    """tests/MRIRecon.cpp's schedule with 5 stages: stop test on, so iteration counts must equal the oracle's."""
    sl = cine[0]
    (x32, r32), (x64, r64) = cine_reference
    dev = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(*dev, capi.nesta_params(**REFERENCE), form=form)
    got = dev[3].get()
    print(f"form {form}: iterations {rep['iters']} (FP32 oracle {r32['iters']}, FP64 {r64['iters']}), "
          f"vs FP32 {rel(got, x32):.2e}, vs FP64 {rel(got, x64):.2e}", flush=True)
    assert rep["iters"] == r32["iters"] == r64["iters"]
    assert rel(got, x32) < 1e-4 and rel(got, x64) < 1e-4


def test_split_precision_workload(ctx, oracle):
    """25 x 16 x 192^2 at acceleration 8: the split form within 1e-3 of FP64 on the fixed schedule (per-iteration
    objective within 1e-4, normU as the oracle's), and its stop-test iteration counts equal the oracle's."""
    sl = synth.make_cine(H=192, W=192, C=16, F=25, accel=8.0, slice_id=3000)
    print("oracle: split workload 25x16x192x192, FP64, fixed and stop-test schedules", flush=True)
    x64, r64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**FIXED), double=True)
    xs64, rs64 = oracle.nesta(sl.y, sl.S, sl.M, oracle.params(**REFERENCE), double=True)
    dev = upload(ctx, sl.y, sl.S, sl.M)
    rep = ctx.nesta(*dev, capi.nesta_params(**FIXED), form=capi.OCL_FORM_SPLIT)
    xs = dev[3].get()
    ctx.nesta(*dev, capi.nesta_params(**FIXED), form=capi.OCL_FORM_FUSED)
    xf = dev[3].get()
    f_err = np.max(np.abs(np.asarray(rep["trace_f"]) - r64["trace_f"]) / np.abs(r64["trace_f"]))
    print(f"split: vs FP64 {rel(xs, x64):.2e}, vs the FP32 form {rel(xs, xf):.2e}, FP32 form vs FP64 "
          f"{rel(xf, x64):.2e}, per-iteration f {f_err:.2e}", flush=True)
    assert rep["iters"] == [20] * 5
    assert abs(rep["normU"] - r64["normU"]) < 2e-6 * r64["normU"]
    assert rel(xs, x64) < 1e-3 and rel(xs, xf) < 1e-3 and rel(xf, x64) < 1e-4
    assert f_err < 1e-4
    rep = ctx.nesta(*dev, capi.nesta_params(**REFERENCE), form=capi.OCL_FORM_SPLIT)
    print(f"split stop test: iterations {rep['iters']} (FP64 oracle {rs64['iters']}), vs FP64 "
          f"{rel(dev[3].get(), xs64):.2e}", flush=True)
    assert rep["iters"] == rs64["iters"]
    assert rel(dev[3].get(), xs64) < 1e-3


def _device_slices(ctx, sizes, id0):
    dev = []
    for i, (h, w, c, f) in enumerate(sizes):
        x, S, M = synth.make_image_parts(h, w, c, f, accel=6.0, slice_id=id0 + i)
        y, dS, dM = synth.device_kspace(ctx, x, S, M, seed=synth.BASE_SEED + id0 + i)
        dev.append((y, dS, dM, ctx.empty(x.shape)))
    ctx.sync()
    return dev


def _solo(ctx, d, prm):
    y, S, M, _ = d
    x = ctx.empty((y.shape[0], y.shape[2], y.shape[3]))
    rep = ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_FUSED)
    return x.get(), rep


def test_uneven_batch_of_64(oracle):
    """The uneven 64-slice table with autotuned tile shapes: every slice bit-equal to its solo run on the same
    context, and the smallest, the largest and the cheapest slice with more than 16 column roles (H = 288, 320)
    within 1e-3 of FP64."""
    c = capi.Context(0)
    try:
        c.set_autotune(True)
        sizes = synth.batch_sizes(64)
        print("uneven batch: generating 64 slices", flush=True)
        dev = _device_slices(c, sizes, 1000)
        prm = capi.nesta_params(maxIntIter=2, maxIter=5, stoptest=0)
        b = c.batch(dev, prm)
        b.setup()
        info = b.info()
        b.iterate(10)
        b.finalize()
        c.sync()
        print("uneven batch", info, flush=True)
        assert info["launch_groups"] >= 2
        out = [d[3].get() for d in dev]
        for i, d in enumerate(dev):
            xs, rs = _solo(c, d, prm)
            assert rs["iters"] == b.report(i)["iters"] == [5, 5]
            assert np.array_equal(out[i], xs), (i, sizes[i])
        work = [h * w * cc * f for (h, w, cc, f) in sizes]
        big_p = [i for i in range(64) if sizes[i][0] >= 288]
        picks = sorted({int(np.argmin(work)), int(np.argmax(work)), min(big_p, key=lambda i: work[i])})
        for i in picks:
            y, S, M = dev[i][0].get(), dev[i][1].get(), dev[i][2].get()
            print("oracle FP64: uneven slice", i, sizes[i], flush=True)
            x64, _ = oracle.nesta(y, S, M, oracle.params(maxIntIter=2, maxIter=5, stoptest=0), double=True)
            e = rel(out[i], x64)
            print(f"uneven slice {i} {sizes[i]}: vs FP64 {e:.2e}", flush=True)
            assert e < 1e-3
        b.close()
    finally:
        c.close()


def test_mixed_table_with_a_launch_group(ctx, oracle):
    """Seven 20 x 8 x 256^2 slices (320 column-pass workgroups each, 2240 together: a launch group of their own) and
    three smaller slices of other sizes (the shared launch of the all-sizes kernels), with the stop test on: every
    slice's iteration counts and result equal its solo run, and two slices are within 1e-4 of FP64."""
    sizes = [(256, 256, 8, 20)] * 7 + [(160, 160, 6, 9), (192, 192, 4, 12), (128, 128, 5, 10)]
    print("mixed table: generating 10 slices", flush=True)
    dev = _device_slices(ctx, sizes, 2000)
    prm = capi.nesta_params(maxIntIter=3, maxIter=12, stoptest=1, miniter=5)
    b = ctx.batch(dev, prm)
    b.setup()
    info = b.info()
    print("mixed table", info, flush=True)
    assert info["launch_groups"] == 2
    b.run()
    out = [d[3].get() for d in dev]
    reps = [b.report(i) for i in range(len(dev))]
    for i, d in enumerate(dev):
        xs, rs = _solo(ctx, d, prm)
        print(i, sizes[i], reps[i]["iters"], rs["iters"], flush=True)
        assert reps[i]["iters"] == rs["iters"]
        assert np.array_equal(out[i], xs)
    for i in (0, 8):
        y, S, M = dev[i][0].get(), dev[i][1].get(), dev[i][2].get()
        x64, r64 = oracle.nesta(y, S, M, oracle.params(maxIntIter=3, maxIter=12, stoptest=1, miniter=5), double=True)
        print(f"mixed slice {i}: vs FP64 {rel(out[i], x64):.2e}, iterations {reps[i]['iters']} / {r64['iters']}",
              flush=True)
        assert rel(out[i], x64) < 1e-4
        assert reps[i]["iters"] == r64["iters"]
    b.close()


def test_every_size_in_a_launch_of_its_own():
    """With the launch-group threshold at one workgroup (ocl_ctx_set_group_threshold), each of three sizes runs its
    own specialized kernels in a launch of its own: every slice's iteration counts and result equal its solo run."""
    sizes = [(160, 160, 6, 9), (192, 192, 4, 12), (128, 128, 5, 10)]
    c = capi.Context(0)
    try:
        with pytest.raises(capi.OclError):
            c.set_group_threshold(0)
        c.set_group_threshold(1)
        print("per-size table: generating 3 slices", flush=True)
        dev = _device_slices(c, sizes, 2700)
        prm = capi.nesta_params(maxIntIter=3, maxIter=12, stoptest=1, miniter=5)
        b = c.batch(dev, prm)
        b.setup()
        info = b.info()
        print("per-size table", info, flush=True)
        assert info["launch_groups"] == 3
        b.run()
        out = [d[3].get() for d in dev]
        reps = [b.report(i) for i in range(len(dev))]
        b.close()
        for i, d in enumerate(dev):
            xs, rs = _solo(c, d, prm)
            print(i, sizes[i], reps[i]["iters"], rs["iters"], flush=True)
            assert reps[i]["iters"] == rs["iters"]
            assert np.array_equal(out[i], xs)
    finally:
        c.close()


def test_one_size_longest_first(ctx):
    """One size with 3, 9 and 6 coils: the table runs as one launch group whose workgroups go longest first (9, 6,
    3 coils), with the stop test on; every slice's iteration counts and result equal its solo run."""
    sizes = [(128, 128, 3, 8), (128, 128, 9, 6), (128, 128, 6, 10)]
    print("one-size table: generating 3 slices", flush=True)
    dev = _device_slices(ctx, sizes, 2500)
    prm = capi.nesta_params(maxIntIter=3, maxIter=12, stoptest=1, miniter=5)
    b = ctx.batch(dev, prm)
    b.setup()
    info = b.info()
    print("one-size table", info, flush=True)
    assert info["launch_groups"] == 1
    b.run()
    out = [d[3].get() for d in dev]
    reps = [b.report(i) for i in range(len(dev))]
    b.close()
    for i, d in enumerate(dev):
        xs, rs = _solo(ctx, d, prm)
        print(i, sizes[i], reps[i]["iters"], rs["iters"], flush=True)
        assert reps[i]["iters"] == rs["iters"]
        assert np.array_equal(out[i], xs)


def test_served_trace_equals_solo(ctx):
    """100 requests over the cine and uneven sizes: five at once, then Poisson arrivals at 50 per second, each with
    5 stages x U{10..30} iterations and its own output buffer.  A 384 MiB budget makes requests wait for memory.
    Every served result is bit-equal to its solo run."""
    import time
    cine_parts = [(256, 256, 8, 20)] * 4
    pool = _device_slices(ctx, cine_parts, 0) + _device_slices(ctx, synth.batch_sizes(64)[:8], 1000)
    rng = np.random.default_rng(5)
    n = 100
    arrivals = np.concatenate([np.zeros(5), np.cumsum(rng.exponential(1 / 50.0, n - 5))])
    pick = rng.integers(0, len(pool), n)
    iters = rng.integers(10, 31, n)
    outs = [ctx.empty((pool[i][0].shape[0], pool[i][0].shape[2], pool[i][0].shape[3])) for i in pick]
    ctx.sync()
    budget = 384 << 20
    srv = capi.Server(0, budget_bytes=budget, chunk=5)
    try:
        t0 = time.perf_counter()
        rids = []
        for r in range(n):
            while time.perf_counter() - t0 < arrivals[r]:
                time.sleep(0.001)
            y, S, M, _ = pool[pick[r]]
            rids.append(srv.submit(y, S, M, outs[r], capi.nesta_params(maxIntIter=5, maxIter=int(iters[r]),
                                                                         stoptest=0), want_report=True))
        for rid in rids:
            assert srv.wait(rid, 300.0) == capi.SERVE_DONE
        st = srv.stats()
        print("served", {k: st[k] for k in ("completed", "failed", "peak_used_bytes", "budget_waits")}, flush=True)
        assert st["completed"] == n and st["failed"] == 0
        assert st["peak_used_bytes"] <= budget and st["budget_waits"] >= 1
        lat = [srv.times(rid)[2] - srv.times(rid)[0] for rid in rids]
        print(f"latency p50 {1e3 * np.percentile(lat, 50):.0f} ms, p99 {1e3 * np.percentile(lat, 99):.0f} ms",
              flush=True)
        served_iters = [srv.report(rid)["iters"] for rid in rids]
    finally:
        srv.close()
    for r in range(n):
        xs, rs = _solo(ctx, pool[pick[r]], capi.nesta_params(maxIntIter=5, maxIter=int(iters[r]), stoptest=0))
        assert served_iters[r] == rs["iters"]
        assert np.array_equal(outs[r].get(), xs), r
