"""NESTA-TV throughput on one MI355X (recons/s), beside the plain FP32 form and the CPU restatement.

Headline workload (one "step"): a batch of B=16 independent cine slices, each F=20 frames x C=8 coils x 256x256,
variable-density ky-line mask at acceleration 6 with 16 centre lines, reconstructed with NESTA-TV for 5 continuation
stages x 20 inner iterations (stop test off, fixed work).  The batch is advanced by the fused batched engine (for
line masks the Gram form: one column pass and one scalar step per iteration, replayed from a hipGraph); `value` is
whole-job recons/s with inputs resident in HBM.

Other legs in the same run (each reported under its own key of the JSON line):
  simple_recon  fully-sampled 8-coil 256x256 SimpleMRIRecon (inverse FFT per coil, conj(S) coil combine)
  plain / fused_single   the plain single-instance FP32 form (reference op sequence, host stop test) and the
                fused single-instance form on the headline slice
  uneven_batch  64 slices of uneven size (H=W in 128..320 step 32, C in 4..32, F in 8..30) in one batched
                launch table per iteration, captured in a hipGraph; k-space synthesised on the device; the plain
                form timed over the same 64 slices beside it
  served        Poisson arrivals (50 req/s, 60 s) mixing the sizes above, 5 stages x 10..30 iterations per request;
                each request's inputs are uploaded from page-locked host memory at its arrival (inside its latency)
                and it is admitted under a binding HBM budget (1 GiB) into the continuous-batching server; p50/p99
                latency and recons/s, a 100-request trace whose every result is compared with its solo run, the
                capacity (96 requests at once) and a second trace at 80% of that capacity; the plain form timed on
                the same mix
  motion_compensated  NESTA-TV with argsMC (U = tTV o MC, fused and plain forms) on the headline slice from the motion-free
                result, with a smooth synthetic deformation (up to 1.5 px) over a centred bound box, plus the
                MotionCompensation / AdjointMotionCompensation kernels alone against their HBM roofline
  cpu           the FP32 CPU restatement on the host cores (cpu_baseline, bounded sample) and rel-L2 of the GPU
                results against the FP64 and FP32 CPU oracles on full schedules

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--batch 16] [--serve-seconds 60] [--legs ...] [--no-ref]
    python bench.py --calibrate     (HBM counter calibration for rocprofv3 --pmc runs)
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NESTA-TV recons/sec and p50/p99 latency vs plain FP32, at ≤1e-3 rel-L2 vs FP64"
H = W = 256
C, F, ACCEL = 8, 20, 6.0
STAGES, ITERS = 5, 20
HBM_PEAK_GBS = 8000.0
UNEVEN_SLICES = 64
SERVE_RATE = 50.0
SPLIT_SHAPE = dict(H=192, W=192, C=16, F=25, accel=8.0)   # split-precision workload, per slice
SPLIT_BATCH = 16
MFMA_BF16_PEAK_TFLOPS = 2500.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_nesta_batched_16x20x8x256.json")


def split_mfma_flops(F, C, H, W, nmax, gram=True):
    """MFMA flops the split-precision passes execute per inner iteration of one slice: every 16x16x32 complex block
    is 12 MFMAs of 16*16*32*2 flops.  Per (frame, 16-column tile, coil) an inverse column DFT of (H/16)(kp/32) blocks
    and a forward one of (H/32)(kp/16) blocks (the Gram pass, and pass CA of the row and column passes); the row and
    column passes add pass B: per (frame, coil, 16-row group) a forward and an inverse row DFT of (W/16)(W/32) blocks
    each."""
    blk = 12 * 16 * 16 * 32 * 2
    kp = (nmax + 31) // 32 * 32
    groups = (nmax + 15) // 16
    b = 0 if gram else F * C * groups * 2 * (W // 16) * (W // 32)
    ca = F * (W // 16) * C * ((H // 16) * (kp // 32) + (H // 32) * (kp // 16))
    return blk * (b + ca)


def dense_dft_flops(F, C, H, W):
    """Algorithmic flops of the dense DFTs of one iteration (forward and adjoint, rows and columns, 8 real flops per
    complex multiply-add): 16 F C H W (H + W)."""
    return 16 * F * C * H * W * (H + W)


def b_alg(F, C, H, W):
    """Algorithmic HBM bytes of one NESTA inner iteration of one slice: read y (8FCN), S (8CN), the u8 mask (FN);
    read x_k, W_k, x_ref and write x_{k+1}, W_{k+1} (5 x 8FN)."""
    N = H * W
    return 8 * N * (F * C + C + 5 * F) + F * N


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------------------------------
# host-side input generation (forked workers, before the GPU is touched)
# ---------------------------------------------------------------------------------------------------------------------
def _gen_cine(i):
    from opencliper_amd import synth
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=ACCEL, slice_id=i)
    return sl.y, sl.S, sl.M


def _gen_parts(job):
    from opencliper_amd import synth
    i, (h, w, c, f) = job
    return synth.make_image_parts(h, w, c, f, accel=ACCEL, slice_id=1000 + i)


def _gen_split(i):
    from opencliper_amd import synth
    return synth.make_image_parts(SPLIT_SHAPE["H"], SPLIT_SHAPE["W"], SPLIT_SHAPE["C"], SPLIT_SHAPE["F"],
                                  accel=SPLIT_SHAPE["accel"], slice_id=3000 + i)


def gen_inputs(args, workers):
    from opencliper_amd import synth
    jobs3 = list(enumerate(synth.batch_sizes(UNEVEN_SLICES))) if ("uneven" in args.legs or "served" in args.legs) \
        else []
    if "served" in args.legs and "uneven" not in args.legs:
        jobs3 = jobs3[:12]
    n_split = SPLIT_BATCH if "split" in args.legs else 0
    if workers <= 1:
        # in-process (no fork): for runs under a profiler that has initialised the GPU before the program started
        cine = [_gen_cine(i) for i in range(args.batch)]
        parts = [_gen_parts(j) for j in jobs3]
        split = [_gen_split(i) for i in range(n_split)]
    else:
        with mp.get_context("fork").Pool(max(1, min(workers, args.batch + len(jobs3)))) as pool:
            r2 = pool.map_async(_gen_cine, range(args.batch))
            r3 = pool.map_async(_gen_parts, jobs3)
            r4 = pool.map_async(_gen_split, range(n_split))
            cine, parts, split = r2.get(), r3.get(), r4.get()
    full = synth.make_fully_sampled(H=256, W=256, C=8, F=1, slice_id=5000) if "simple" in args.legs else None
    return cine, parts, full, split


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


# ---------------------------------------------------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------------------------------------------------
def leg_simple(ctx, full, reps=200):
    """SimpleMRIRecon, 8 coils 256x256: device time of the recon per call (the recon works in place on y, so the
    input is restored by a device copy outside the timed region before every call)."""
    y0, S, x = ctx.to_device(full.y), ctx.to_device(full.S), ctx.empty((1, full.H, full.W))
    work = ctx.empty(full.y.shape)
    for _ in range(3):
        ctx.copy(y0, work)
        ctx.simple_mri_recon(work, S, x)
    ms = []
    for _ in range(reps):
        ctx.copy(y0, work)
        t = ctx.timer_begin()
        ctx.simple_mri_recon(work, S, x)
        ms.append(ctx.timer_end(t))
    N = full.H * full.W
    byts = 8 * N * (2 * full.C + 1)
    mean = float(np.mean(ms))
    # the batched entry point: 64 inputs of their own (restored outside the timed region), two launches in all
    nb = 64
    works = [ctx.empty(full.y.shape) for _ in range(nb)]
    Ss = [ctx.to_device(full.S) for _ in range(nb)]
    xs = [ctx.empty((1, full.H, full.W)) for _ in range(nb)]
    items = list(zip(works, Ss, xs))
    bms = []
    for r in range(3 + 20):
        for w in works:
            ctx.copy(y0, w)
        t = ctx.timer_begin()
        ctx.simple_mri_recon_batch(items)
        if r >= 3:
            bms.append(ctx.timer_end(t))
        else:
            ctx.timer_end(t)
    bmean = float(np.mean(bms))
    xb = xs[nb - 1].get()
    batch = {"inputs": nb, "recons_per_s": nb * 1e3 / bmean, "ms_mean": bmean, "ms_p50": pct(bms, 50),
             "ms_p99": pct(bms, 99), "achieved_GBs": nb * byts / (bmean * 1e-3) / 1e9,
             "frac_hbm": nb * byts / (bmean * 1e-3) / 1e9 / HBM_PEAK_GBS,
             "max_abs_diff_vs_single": float(np.max(np.abs(xb - x.get())))}
    return {"recons_per_s": 1e3 / mean, "ms_mean": mean, "ms_p50": pct(ms, 50), "ms_p99": pct(ms, 99),
            "bytes_alg": byts, "achieved_GBs": byts / (mean * 1e-3) / 1e9,
            "frac_hbm": byts / (mean * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "workload": "fully sampled, C=8, 256x256, one frame", "batched": batch}, x.get()


def leg_single(ctx, capi, dev0, prm, form, steps):
    y0, S0, M0, x0 = dev0
    lat = []
    for r in range(1 + steps):
        t = time.perf_counter()
        ctx.nesta(y0, S0, M0, x0, prm, form=form)
        if r:
            lat.append(time.perf_counter() - t)
    return {"recons_per_s": 1.0 / np.mean(lat), "p50_ms": 1e3 * pct(lat, 50), "p99_ms": 1e3 * pct(lat, 99)}, x0.get()


def leg_batched(ctx, dev, prm, steps, warmup, form=None):
    """Whole schedule captured once into a hipGraph; each step = setup (x_ref, normU, mu0, state) + replay +
    finalize for every slice of the batch."""
    from opencliper_amd import capi
    bt = ctx.batch(dev, prm, form=capi.OCL_FORM_FUSED if form is None else form)
    bt.setup()
    bt.capture(STAGES * ITERS)
    ctx.sync()
    for _ in range(warmup):
        bt.setup()
        bt.replay()
        bt.finalize()
        ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        bt.setup()
        bt.replay()
        bt.finalize()
        ctx.sync()
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    B = len(dev)
    res = {"recons_per_s": B * steps / total, "ms_per_step": 1e3 * total / steps,
           "p50_ms": 1e3 * pct(lat, 50), "p99_ms": 1e3 * pct(lat, 99)}
    # per-launch HIP-event timing of the fused iteration on the library's stream
    bt.setup()
    ms_b, ms_ca, ms_tot = bt.time_iters(STAGES * ITERS)
    byts = sum(b_alg(*y.shape) for (y, _, _, _) in dev)
    it_ms = ms_b + ms_ca
    res["iteration"] = {"ms_B_plus_S": ms_b, "ms_CA": ms_ca, "ms_iter": it_ms,
                        "span_ms_per_iter": ms_tot / (STAGES * ITERS), "bytes_alg_per_iter": byts,
                        "achieved_GBs": byts / (it_ms * 1e-3) / 1e9,
                        "frac_hbm": byts / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    res["reports"] = [bt.report(0)]
    bt.close()
    return res


def leg_uneven(ctx, synth, parts, prm, steps, warmup, group_threshold=None):
    """64 uneven slices; the tile shape of each H x W comes from the autotuner (tuned at batch setup, outside the
    timed steps).  group_threshold: column-pass workgroups from which a size gets launches of its own."""
    ctx.set_autotune(True)
    if group_threshold:
        ctx.set_group_threshold(group_threshold)
    dev = []
    for i, (x, S, M) in enumerate(parts):
        y, dS, dM = synth.device_kspace(ctx, x, S, M, seed=synth.BASE_SEED + 1000 + i)
        dev.append((y, dS, dM, ctx.empty(x.shape)))
    ctx.sync()
    t0 = time.perf_counter()
    tiles = {f"{h}x{h}": ctx.autotune_shape(h, h) for h in sorted({x.shape[1] for (x, _, _) in parts})}
    for t in tiles.values():     # ca_fast 3: the Gram form (one column pass per iteration) beat the row + column passes
        t["form"] = "gram" if t["ca_fast"] == 3 else "row and column passes"
    tune_s = time.perf_counter() - t0
    res = leg_batched(ctx, dev, prm, steps, warmup)
    res["autotune"] = {"tune_s": tune_s, "tiles": tiles}
    shapes = [y.shape for (y, _, _, _) in dev]
    res["slices"] = len(dev)
    res["shapes_FCHW"] = shapes
    res["coil_images"] = int(sum(f * c for (f, c, _, _) in shapes))
    return res, dev


def leg_split(ctx, synth, parts, prm, steps, warmup):
    """Split-precision form (BF16-MFMA DFTs, 3-term split) beside the FP32 fused form on the same batch: the Gram pass
    (the default for line masks) and the row and column passes."""
    from opencliper_amd import capi
    dev = []
    for i, (x, S, M) in enumerate(parts):
        y, dS, dM = synth.device_kspace(ctx, x, S, M, seed=synth.BASE_SEED + 3000 + i)
        dev.append((y, dS, dM, ctx.empty(x.shape)))
    ctx.sync()
    res = {"fp32_fused": leg_batched(ctx, dev, prm, steps, warmup, capi.OCL_FORM_FUSED)}
    x32 = dev[0][3].get()
    ctx.set_split_gram(False)
    try:
        res["split_rowcol"] = leg_batched(ctx, dev, prm, steps, warmup, capi.OCL_FORM_SPLIT)
    finally:
        ctx.set_split_gram(True)
    x_rc = dev[0][3].get()
    res["split"] = leg_batched(ctx, dev, prm, steps, warmup, capi.OCL_FORM_SPLIT)
    xs = dev[0][3].get()
    sh = SPLIT_SHAPE
    nmax = int(round(sh["H"] / sh["accel"]))
    dense = len(dev) * dense_dft_flops(sh["F"], sh["C"], sh["H"], sh["W"])
    for key, gram in (("split", True), ("split_rowcol", False)):
        fl = len(dev) * split_mfma_flops(sh["F"], sh["C"], sh["H"], sh["W"], nmax, gram=gram)
        it = res[key]["iteration"]
        it["mfma_flops_per_iter"] = fl
        it["mfma_TFLOPs"] = fl / (it["ms_iter"] * 1e-3) / 1e12
        it["frac_mfma"] = it["mfma_TFLOPs"] / MFMA_BF16_PEAK_TFLOPS
        it["dense_dft_flops_per_iter"] = dense
        it["dense_dft_TFLOPs_equiv"] = dense / (it["ms_iter"] * 1e-3) / 1e12
    host = (dev[0][0].get(), dev[0][1].get(), dev[0][2].get())
    return res, {"x_split": xs, "x_split_rowcol": x_rc, "x_fp32": x32, "slice0": host}


def _serve_trace(ctx, capi, host_pool, arrivals, pick, iters, budget, chunk, check_all=False):
    """Runs one arrival trace against a fresh server.  Each request starts on the host: at its arrival its inputs
    (y, S, M of the pool slice it draws) are uploaded into device buffers of its own (recycled once it is done), then
    it is submitted with an output buffer of its own.  Latency is arrival -> done, so it includes the upload and any
    wait for HBM budget.  check_all: every output buffer is kept and compared afterwards with a solo run."""
    n_req = len(arrivals)
    free_in = [[] for _ in host_pool]
    free_out = [[] for _ in host_pool]

    def take(i):
        y, S, M = host_pool[i]
        if free_in[i]:
            dy, dS, dM = free_in[i].pop()
        else:
            dy, dS, dM = ctx.empty(y.shape), ctx.empty(S.shape), ctx.empty(M.shape, np.uint8)
        dy.set(y)
        dS.set(S)
        dM.set(M)
        xo = free_out[i].pop() if (free_out[i] and not check_all) else ctx.empty((y.shape[0], y.shape[2], y.shape[3]))
        return dy, dS, dM, xo

    srv = capi.Server(0, budget_bytes=budget, chunk=chunk)
    inflight, meta, arrive_py, submit_py, upload_s = {}, {}, {}, {}, []
    t0 = time.perf_counter()
    t_log = t0
    try:
        for r in range(n_req):
            if time.perf_counter() - t_log > 20.0:
                t_log = time.perf_counter()
                log(f"served trace: {t_log - t0:.0f} s, {r} of {n_req} submitted")
            while True:
                for rid in [k for k in inflight if srv.poll(k)[0] == capi.SERVE_DONE]:
                    i, bufs = inflight.pop(rid)
                    free_in[i].append(bufs[:3])
                    if not check_all:
                        free_out[i].append(bufs[3])
                now = time.perf_counter() - t0
                if now >= arrivals[r]:
                    break
                time.sleep(min(0.002, arrivals[r] - now))
            i = int(pick[r])
            tu = time.perf_counter()
            bufs = take(i)
            upload_s.append(time.perf_counter() - tu)
            prm = capi.nesta_params(maxIntIter=STAGES, maxIter=int(iters[r]), stoptest=0)
            ts = time.perf_counter() - t0
            rid = srv.submit(*bufs, prm)
            inflight[rid] = (i, bufs)
            meta[rid] = (i, int(iters[r]), bufs[3])
            arrive_py[rid] = float(arrivals[r])
            submit_py[rid] = ts
        for rid in list(inflight):
            srv.wait(rid, 600.0)
        stats = srv.stats()
        times = {rid: srv.times(rid) for rid in meta}
    finally:
        srv.close()
    # arrival -> submit on the host clock (upload and the submit loop), submit -> done on the server's clock
    lat = [(submit_py[k] - arrive_py[k]) + (times[k][2] - times[k][0]) for k in times]
    qwait = [times[k][1] - times[k][0] for k in times]
    t_first = min(v[0] for v in times.values())
    t_last = max(v[2] for v in times.values())
    res = {"requests": len(times), "recons_per_s": len(times) / (t_last - t_first),
           "p50_ms": 1e3 * pct(lat, 50), "p99_ms": 1e3 * pct(lat, 99), "max_ms": 1e3 * max(lat),
           "upload_ms_mean": 1e3 * float(np.mean(upload_s)), "queue_wait_p50_ms": 1e3 * pct(qwait, 50),
           "queue_wait_p99_ms": 1e3 * pct(qwait, 99), "slice_iterations": sum(STAGES * meta[k][1] for k in times),
           "budget_bytes": budget, "peak_used_bytes": stats["peak_used_bytes"],
           "budget_waits": stats["budget_waits"], "launches": stats["launches"], "busy_s": stats["busy_s"],
           "buffer_cache_hits": stats["cache_hits"], "buffer_cache_misses": stats["cache_misses"],
           "admit_s": stats["admit_s"], "retire_s": stats["retire_s"],
           "failed": stats["failed"]}
    return res, meta


def leg_served(ctx, capi, host_pool, seconds, rate, budget, chunk, seed=7, n_check=None):
    """Poisson trace at `rate` requests/s for `seconds`, 5 stages x U{10..30} iterations per request, sizes drawn
    from the pool (the headline cine slices and the uneven sizes).  With n_check, that many requests (spread over the
    trace) keep their output buffers and are compared with solo runs afterwards."""
    rng = np.random.default_rng(seed)
    n_req = max(1, int(round(rate * seconds)))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    pick = rng.integers(0, len(host_pool), n_req)
    iters = rng.integers(10, 31, n_req)
    res, meta = _serve_trace(ctx, capi, host_pool, arrivals, pick, iters, budget, chunk)
    res.update({"offered_rate_per_s": rate, "trace_s": float(arrivals[-1]), "pool_slices": len(host_pool)})
    if n_check:
        # a second, shorter trace at the same rate whose every output is kept: each served result against its solo run
        m = min(n_req, n_check)
        chk, cmeta = _serve_trace(ctx, capi, host_pool, arrivals[:m], pick[:m], iters[:m], budget, chunk,
                                  check_all=True)
        eq = []
        dev = {}
        for rid, (i, it, xo) in cmeta.items():
            if i not in dev:
                y, S, M = host_pool[i]
                dev[i] = (ctx.to_device(y), ctx.to_device(S), ctx.to_device(M))
            y, S, M = dev[i]
            xs = ctx.empty((y.shape[0], y.shape[2], y.shape[3]))
            ctx.nesta(y, S, M, xs, capi.nesta_params(maxIntIter=STAGES, maxIter=it, stoptest=0),
                      form=capi.OCL_FORM_FUSED)
            a, b = xo.get(), xs.get()
            eq.append(float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())))
        res["served_vs_solo"] = {"requests": len(eq), "rel_l2_max": max(eq), "bit_equal": int(sum(e == 0.0 for e in eq)),
                                 "budget_waits": chk["budget_waits"]}
    return res


def leg_served_burst(ctx, capi, host_pool, n_req, budget, chunk, seed=11):
    """Capacity of the server: n_req requests of the same mix arriving at once (each uploaded from the host)."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(host_pool), n_req)
    iters = rng.integers(10, 31, n_req)
    res, _ = _serve_trace(ctx, capi, host_pool, np.zeros(n_req), pick, iters, budget, chunk)
    return res


def leg_plain_mix(ctx, capi, host_pool, n_req, seed=7):
    """The plain single-instance FP32 form over the served request mix (same pool and iteration distribution as the
    trace): recons/s with inputs resident, one request after another."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(host_pool), n_req)
    iters = rng.integers(10, 31, n_req)
    dev = {}
    for i in sorted(set(int(v) for v in pick)):
        y, S, M = host_pool[i]
        dev[i] = (ctx.to_device(y), ctx.to_device(S), ctx.to_device(M),
                  ctx.empty((y.shape[0], y.shape[2], y.shape[3])))
    ctx.sync()
    t = time.perf_counter()
    for r in range(n_req):
        ctx.nesta(*dev[int(pick[r])], capi.nesta_params(maxIntIter=STAGES, maxIter=int(iters[r]), stoptest=0),
                  form=capi.OCL_FORM_PLAIN)
    total = time.perf_counter() - t
    return {"requests": n_req, "recons_per_s": n_req / total}


def leg_plain_uneven(ctx, capi, dev3, prm):
    """The plain single-instance FP32 form over the 64 uneven slices, one after another (inputs resident)."""
    t = time.perf_counter()
    for (y, S, M, x) in dev3:
        ctx.nesta(y, S, M, x, prm, form=capi.OCL_FORM_PLAIN)
    total = time.perf_counter() - t
    return {"slices": len(dev3), "recons_per_s": len(dev3) / total, "s_total": total}


def mc_bytes(F, H, W, box_pixels, adjoint):
    """Algorithmic HBM bytes of one MotionCompensation (copy 16 B per pixel-frame, then per box pixel-frame the
    deformed position 8 B, one input sample 8 B and the output 8 B, plus the 8-B mesh per box pixel) or
    AdjointMotionCompensation (copy with the box cleared 16 B per pixel-frame, then per box pixel-frame the position
    8 B, the input 8 B and the read-modify-write of the scattered output 16 B, plus the mesh)."""
    per_box = 32 if adjoint else 24
    return 16 * F * H * W + per_box * F * box_pixels + 8 * box_pixels


def leg_mc(ctx, synth, dev0, x_ref, prm, steps, reps=20):
    y0, S0, M0, _ = dev0
    Fh, Hh, Wh = x_ref.shape
    xn, mesh, bb = synth.motion_fields(Hh, Wh, Fh, amp=1.5, slice_id=77)
    dxn, dmesh = ctx.to_device(xn), ctx.to_device(mesh)
    mc = ctx.mc_args(dxn, dmesh, bb)
    m0, m1 = Hh // 40, Wh // 40
    nr, nc = int(bb[3]) - int(bb[1]) + 2 * m0 + 3, int(bb[2]) - int(bb[0]) + 2 * m1 + 3
    xin, out = ctx.to_device(x_ref), ctx.empty(x_ref.shape)
    kern = {}
    for adjoint in (False, True):
        for _ in range(3):
            ctx.motion_compensation(xin, out, mc, adjoint=adjoint)
        ms = []
        for _ in range(reps):
            t = ctx.timer_begin()
            ctx.motion_compensation(xin, out, mc, adjoint=adjoint)
            ms.append(ctx.timer_end(t))
        byts = mc_bytes(Fh, Hh, Wh, nr * nc, adjoint)
        mean = float(np.mean(ms))
        kern["adjoint" if adjoint else "forward"] = {
            "ms_mean": mean, "bytes_alg": byts, "achieved_GBs": byts / (mean * 1e-3) / 1e9,
            "frac_hbm": byts / (mean * 1e-3) / 1e9 / HBM_PEAK_GBS}
    from opencliper_amd import capi
    x = ctx.empty(x_ref.shape)
    forms, xs = {}, {}
    for name, form, reps in (("plain", capi.OCL_FORM_PLAIN, steps), ("fused", capi.OCL_FORM_FUSED, 3 * steps)):
        lat, rep = [], None
        for r in range(1 + reps):
            x.set(x_ref)
            t = time.perf_counter()
            rep = ctx.nesta_mc(y0, S0, M0, x, mc, prm, form=form)
            if r:
                lat.append(time.perf_counter() - t)
        forms[name] = {"recons_per_s": 1.0 / np.mean(lat), "p50_ms": 1e3 * pct(lat, 50),
                       "p99_ms": 1e3 * pct(lat, 99), "normU": rep["normU"]}
        xs[name] = x.get()
    return {"recons_per_s": forms["fused"]["recons_per_s"], "forms": forms, "box_rows_cols": [nr, nc],
            "bound_box": [int(v) for v in bb], "kernels": kern,
            "workload": f"headline slice F={Fh} C={C} {Hh}x{Wh}, {STAGES} stages x {ITERS} iterations, from the "
                        f"motion-free reconstruction; deformation up to 1.5 px; fused form (plain beside it)"}, \
        xs, (xn, mesh, bb)


def gwr_bytes(F, L, Cp):
    """Algorithmic HBM bytes of one registration iteration: the eval pass reads the 16-B {I0, d1, d2} sample of every
    box pixel-frame and writes its two mesh coordinates and two gradient weights (4 B each), the update pass reads the
    weights again; the coefficients (2 F Cp^2 floats) are read by both passes and written once."""
    return F * L * L * (16 + 8 + 8 + 8) + 3 * 2 * F * Cp * Cp * 4


def leg_registration(ctx, capi, x, reps=5):
    """GroupwiseRegistration of the headline slice's plain reconstruction with tests/MRIRecon.cpp's parameters, as
    the reference's driver runs it before motion-compensated NESTA; then 100 fixed iterations (no stop rule) and one
    iteration, whose difference gives the time per iteration."""
    F, H, W = x.shape
    dx = ctx.to_device(x.astype(np.complex64))
    out = {}
    for name, prm in (("mrirecon", capi.gwr_params()), ("fixed100", capi.gwr_params(nmax=100, et=0.0, eh=0.0)),
                      ("fixed1", capi.gwr_params(nmax=1, et=0.0, eh=0.0))):
        reg = ctx.registration(F, H, W, prm)
        T = ctx.empty((2, F, reg.Cp, reg.Cp), np.float32)
        xn = ctx.empty((2, F, H, W), np.float32)
        rep = reg.run(dx, T, xn)
        lat = []
        for _ in range(reps):
            t = time.perf_counter()
            rep = reg.run(dx, T, xn)
            lat.append(time.perf_counter() - t)
        out[name] = {"ms": 1e3 * float(np.median(lat)), "iters": rep["iters"], "H0": rep["H0"],
                     "H_final": float(rep["H"][rep["iters"]])}
        geo = (reg.Cp, reg.box)
        reg.close()
    us_it = 1e3 * (out["fixed100"]["ms"] - out["fixed1"]["ms"]) / 99
    byts = gwr_bytes(F, geo[1], geo[0])
    m = out["mrirecon"]
    return {"workload": f"registration of the plain reconstruction F={F} {H}x{W}: circular ROI radius 40, cubic "
                        f"B-splines Dp 4 ({geo[0]}x{geo[0]} control points), W 1, flagW, nmax 100, et 0.01, "
                        f"eh 0.005, lambda (0, 0.005, 0, 0.5) (tests/MRIRecon.cpp:122-130)",
            "ms": m["ms"], "iters": m["iters"], "H0": m["H0"], "H_final": m["H_final"],
            "registrations_per_s": 1e3 / m["ms"], "us_per_iteration": us_it,
            "roofline": {"bound": "hbm", "achieved": byts / (us_it * 1e-6) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": byts / (us_it * 1e-6) / 1e9 / HBM_PEAK_GBS,
                         "bytes_alg_per_iter": byts,
                         "kernel": "k_gwr_grad_update + k_gwr_eval + k_gwr_step (one optimizer iteration, three dependent launches; "
                                   "wall clock of 100 minus 1 fixed iterations)"}}, x


def gpu_legs(args, cine, parts, full, split):
    from opencliper_amd import capi, synth
    ctx = capi.Context(0)
    out = {"device": ctx.info()["name"], "x": {}}
    prm = capi.nesta_params(maxIntIter=STAGES, maxIter=ITERS, stoptest=0)
    try:
        if "simple" in args.legs:
            out["simple_recon"], out["x"]["simple"] = leg_simple(ctx, full)
            log(f"simple_recon: {out['simple_recon']['recons_per_s']:.0f} recons/s, batched "
                f"{out['simple_recon']['batched']['recons_per_s']:.0f} recons/s")
        dev = [(ctx.to_device(y), ctx.to_device(S), ctx.to_device(M), ctx.empty((F, H, W))) for (y, S, M) in cine]
        out["plain"], out["x"]["plain"] = leg_single(ctx, capi, dev[0], prm, capi.OCL_FORM_PLAIN, args.plain_steps)
        log(f"plain: {out['plain']['recons_per_s']:.2f} recons/s")
        if "mc" in args.legs:
            out["mc"], out["x"]["mc"], out["x"]["mc_fields"] = leg_mc(ctx, synth, dev[0], out["x"]["plain"], prm,
                                                                      args.plain_steps)
            k = out["mc"]["kernels"]
            log(f"motion-compensated: fused {out['mc']['forms']['fused']['recons_per_s']:.2f} recons/s, plain "
                f"{out['mc']['forms']['plain']['recons_per_s']:.2f}; MC kernel {k['forward']['achieved_GBs']:.0f} "
                f"GB/s, adjoint {k['adjoint']['achieved_GBs']:.0f} GB/s")
        if "reg" in args.legs:
            log("registration leg")
            out["registration"], out["x"]["reg_input"] = leg_registration(ctx, capi, out["x"]["plain"])
        out["fused_single"], out["x"]["fused"] = leg_single(ctx, capi, dev[0], prm, capi.OCL_FORM_FUSED, args.steps)
        log(f"fused single: {out['fused_single']['recons_per_s']:.2f} recons/s")
        out["batched"] = leg_batched(ctx, dev, prm, args.steps, args.warmup)
        out["x"]["batched0"] = dev[0][3].get()
        log(f"batched: {out['batched']['recons_per_s']:.1f} recons/s, iteration {out['batched']['iteration']}")
        if "split" in args.legs:
            out["split"], out["x"]["split"] = leg_split(ctx, synth, split, prm, args.steps, args.warmup)
            sp = out["split"]
            log(f"split: {sp['split']['recons_per_s']:.1f} recons/s, row and column passes "
                f"{sp['split_rowcol']['recons_per_s']:.1f} (FP32 fused on the same batch "
                f"{sp['fp32_fused']['recons_per_s']:.1f}), iteration {sp['split']['iteration']}")
        dev3 = []
        if "uneven" in args.legs:
            out["uneven_batch"], dev3 = leg_uneven(ctx, synth, parts, prm, args.steps, args.warmup,
                                                   args.group_threshold)
            small = min(range(len(dev3)), key=lambda i: b_alg(*dev3[i][0].shape))
            out["x"]["uneven_small"] = (small, dev3[small][0].get(), dev3[small][1].get(), dev3[small][2].get(),
                                        dev3[small][3].get())
            u = out["uneven_batch"]
            log(f"uneven batch: {u['recons_per_s']:.1f} recons/s, iteration {u['iteration']}")
        if "uneven" in args.legs and dev3:
            out["plain_uneven"] = leg_plain_uneven(ctx, capi, dev3, prm)
            log(f"plain over the uneven batch: {out['plain_uneven']['recons_per_s']:.2f} recons/s")
        if "served" in args.legs:
            if not dev3:
                for i, (x, S, M) in enumerate(parts):
                    y, dS, dM = synth.device_kspace(ctx, x, S, M, seed=synth.BASE_SEED + 1000 + i)
                    dev3.append((y, dS, dM, None))
            # request payloads wait in page-locked host memory, as a server's receive buffers would: each upload is
            # one DMA transfer (from pageable memory the driver stages it, at about a fifth of the rate)
            host_pool = [(ctx.pinned(y.get()), ctx.pinned(S.get()), ctx.pinned(M.get()))
                         for (y, S, M, _) in dev[:4] + dev3[:12]]
            out["plain_served_mix"] = leg_plain_mix(ctx, capi, host_pool, 16)
            log(f"plain over the served mix: {out['plain_served_mix']['recons_per_s']:.2f} recons/s")
            budget = int(args.serve_budget * (1 << 30))
            out["served"] = leg_served(ctx, capi, host_pool, args.serve_seconds, SERVE_RATE, budget,
                                       args.serve_chunk, n_check=args.serve_check)
            s = out["served"]
            log(f"served: {s['requests']} requests, {s['recons_per_s']:.1f} recons/s, p50 {s['p50_ms']:.0f} ms, "
                f"p99 {s['p99_ms']:.0f} ms, budget waits {s['budget_waits']}")
            if args.serve_burst > 0:
                out["served_burst"] = leg_served_burst(ctx, capi, host_pool, args.serve_burst, budget,
                                                       args.serve_chunk)
                s = out["served_burst"]
                log(f"served burst: {s['requests']} requests, {s['recons_per_s']:.1f} recons/s")
                if args.serve_burst_budget > 0:
                    # the same burst under a budget that holds several times more requests at once (it still binds:
                    # the 96 requests need more), so that capacity is set by the device rather than by admission
                    b2 = int(args.serve_burst_budget * (1 << 30))
                    out["served_burst_large_budget"] = leg_served_burst(ctx, capi, host_pool, args.serve_burst, b2,
                                                                        args.serve_chunk)
                    s2 = out["served_burst_large_budget"]
                    if out.get("plain_served_mix"):
                        s2["speedup_vs_plain"] = s2["recons_per_s"] / out["plain_served_mix"]["recons_per_s"]
                    log(f"served burst, {args.serve_burst_budget:g} GiB budget: {s2['recons_per_s']:.1f} recons/s, "
                        f"peak {s2['peak_used_bytes'] / 2**30:.2f} GiB, budget waits {s2['budget_waits']}")
                if args.serve_load_seconds > 0:
                    rate = 0.8 * s["recons_per_s"]
                    out["served_80pct"] = leg_served(ctx, capi, host_pool, args.serve_load_seconds, rate, budget,
                                                     args.serve_chunk, seed=13)
                    s = out["served_80pct"]
                    log(f"served at 80% of capacity ({rate:.1f} req/s): {s['recons_per_s']:.1f} recons/s, "
                        f"p50 {s['p50_ms']:.0f} ms, p99 {s['p99_ms']:.0f} ms")
    finally:
        ctx.close()
    return out


# ---------------------------------------------------------------------------------------------------------------------
# CPU leg: oracle (the FP32 restatement as the baseline, FP64 and FP32 as the checkers)
# ---------------------------------------------------------------------------------------------------------------------
def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "os_cpu_count": os.cpu_count()}


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def cpu_leg(args, slice0, full, g):
    """FP32 restatement timed on the host: config 2's full schedule (setup + 5 stages x 20 iterations) on all threads,
    the median of 3 runs after a one-iteration warm-up; the 1-thread figure extrapolated from single runs of 1 and 2
    iterations (a full 1-thread schedule would take minutes); config 1 as the median of 5 (and of 3 on one thread)."""
    from oracle import oracle as O
    y, S, M = slice0
    threads = O.num_threads()
    res = {}
    run = lambda n, k=1: O.nesta(y, S, M, O.params(maxIntIter=k, maxIter=n, stoptest=0), double=False)
    run(1)                                                                    # warm-up
    t_recon = _median_time(lambda: run(ITERS, STAGES), 3)
    O.set_threads(1)
    try:
        t1a = _median_time(lambda: run(1), 1)
        t1b = _median_time(lambda: run(2), 1)
    finally:
        O.set_threads(threads)
    t1_iter = max(1e-9, t1b - t1a)
    t1_recon = max(0.0, t1a - t1_iter) + STAGES * ITERS * t1_iter
    res["cpu_baseline"] = {"value": 1.0 / t_recon, "unit": "recons/s", "cores": threads, "kind": "port",
                           "sample": f"FP32 restatement (oracle/nesta_oracle.cpp, OpenMP, {threads} threads), one "
                                     f"20x8x256x256 slice, the full schedule (setup + {STAGES} stages x {ITERS} "
                                     f"iterations): median of 3 runs after a one-iteration warm-up, "
                                     f"{t_recon:.2f} s per reconstruction",
                           "one_thread_recons_per_s": 1.0 / t1_recon,
                           "one_thread_sample": f"extrapolated from single runs of 1 and 2 iterations on 1 thread "
                                                f"({t1_iter:.3f} s/iteration)",
                           "host": host_info()}
    gx = g["x"]
    if full is not None and "simple" in gx:
        O.simple_recon(full.y, full.S, double=False)
        t_med = _median_time(lambda: O.simple_recon(full.y, full.S, double=False), 5)
        O.set_threads(1)
        try:
            t_one = _median_time(lambda: O.simple_recon(full.y, full.S, double=False), 3)
        finally:
            O.set_threads(threads)
        x64 = O.simple_recon(full.y, full.S, double=True)
        res["simple_recon"] = {"cpu_recons_per_s": 1.0 / t_med, "cpu_one_thread_recons_per_s": 1.0 / t_one,
                               "cores": threads,
                               "rel_l2_vs_fp64": O.rel_l2(gx["simple"], x64),
                               "rel_l2_vs_truth": O.rel_l2(gx["simple"], full.x_true)}
    if "reg_input" in gx:
        from oracle import gwr_oracle as G
        xr = gx["reg_input"]
        t = time.perf_counter()
        R = G.Registration(xr.shape[1], xr.shape[0])
        t_init = time.perf_counter() - t
        t = time.perf_counter()
        R.run(xr, nmax=1)
        t_it = time.perf_counter() - t
        iters = g["registration"]["iters"]
        res["registration"] = {"cpu_s_init": t_init, "cpu_s_first_iteration": t_it,
                               "cpu_s_per_registration_extrapolated": t_init + iters * t_it, "cores": 1,
                               "kind": "port",
                               "sample": "oracle/gwr_oracle.py (numpy, kernel by kernel, FP64) on the same "
                                         "reconstruction: init + creabspline and one optimizer iteration, "
                                         f"extrapolated to the GPU run's {iters} iterations"}
    if not args.no_ref:
        prm = O.params(maxIntIter=STAGES, maxIter=ITERS, stoptest=0)
        x64, _ = O.nesta(y, S, M, prm, double=True)
        x32, _ = O.nesta(y, S, M, prm, double=False)
        rl = {form: {"vs_fp64": O.rel_l2(gx[form], x64), "vs_fp32_restatement": O.rel_l2(gx[form], x32)}
              for form in ("plain", "fused", "batched0")}
        rl["fp32_restatement_vs_fp64"] = O.rel_l2(x32, x64)
        if "split" in gx:
            sx = gx["split"]
            yy, SS, MM = sx["slice0"]
            ref, r64 = O.nesta(yy, SS, MM, prm, double=True)
            f64 = np.asarray(r64["trace_f"])
            fs = np.asarray(g["split"]["split"]["reports"][0]["trace_f"])
            f32 = np.asarray(g["split"]["fp32_fused"]["reports"][0]["trace_f"])
            rl["split"] = {"shape_FCHW": list(yy.shape), "vs_fp64": O.rel_l2(sx["x_split"], ref),
                           "vs_fp32_form": O.rel_l2(sx["x_split"], sx["x_fp32"]),
                           "rowcol_vs_fp64": O.rel_l2(sx["x_split_rowcol"], ref),
                           "fp32_form_vs_fp64": O.rel_l2(sx["x_fp32"], ref),
                           "per_iter_f_relerr_vs_fp64_max": float(np.max(np.abs(fs - f64) / np.abs(f64))),
                           "per_iter_f_relerr_vs_fp32_max": float(np.max(np.abs(fs - f32) / np.abs(f32))),
                           "per_iter_f_relerr_vs_fp64": [float(v) for v in np.abs(fs - f64) / np.abs(f64)]}
        if "mc" in gx:
            xn, mesh, bb = gx["mc_fields"]
            ref, _ = O.nesta_mc(y, S, M, gx["plain"], xn, mesh, bb, prm, double=True)
            rl["motion_compensated"] = {"fused_vs_fp64": O.rel_l2(gx["mc"]["fused"], ref),
                                        "plain_vs_fp64": O.rel_l2(gx["mc"]["plain"], ref)}
        if "uneven_small" in gx:
            i, yy, SS, MM, xx = gx["uneven_small"]
            ref, _ = O.nesta(yy, SS, MM, prm, double=True)
            rl["uneven_batch_smallest"] = {"slice": i, "shape_FCHW": list(yy.shape), "vs_fp64": O.rel_l2(xx, ref)}
        res["rel_l2"] = rl
    return res


CALIB_BYTES = 256 << 20


def calibrate():
    """HBM byte-counter calibration for rocprofv3 --pmc runs: the library's copy kernel over 256 MiB, 5 dispatches.
    Each dispatch reads and writes 256 MiB, so its FETCH_SIZE / WRITE_SIZE (KiB per dispatch) give the factors that
    turn the same counters of the library's other kernels into bytes."""
    from opencliper_amd import capi
    ctx = capi.Context(0)
    try:
        n = CALIB_BYTES // 8
        a = ctx.zeros((n,), np.complex64)
        b = ctx.empty((n,), np.complex64)
        for _ in range(5):
            ctx.copy(a, b)
        ctx.sync()
        print(f"copy kernel: {CALIB_BYTES} bytes read + {CALIB_BYTES} bytes written per dispatch, 5 dispatches")
    finally:
        ctx.close()


def load_traffic():
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--legs", default="simple,mc,reg,split,uneven,served",
                    help="comma list of extra legs beside the headline batch: simple, mc, split, uneven, served")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the Poisson trace (300 = 5 min)")
    ap.add_argument("--serve-budget", type=float, default=1.0,
                    help="server HBM budget in GiB (the default binds: requests wait for memory)")
    ap.add_argument("--serve-chunk", type=int, default=5, help="server iterations between admission points")
    ap.add_argument("--serve-burst", type=int, default=96, help="requests submitted at once for server capacity")
    ap.add_argument("--group-threshold", type=int, default=None,
                    help="uneven batch: column-pass workgroups from which a size gets launches of its own")
    ap.add_argument("--serve-burst-budget", type=float, default=0.0,
                    help="HBM budget (GiB) of a second capacity burst (0: skip)")
    ap.add_argument("--serve-load-seconds", type=float, default=30.0,
                    help="length of a second Poisson trace at 80%% of the measured capacity (0: skip)")
    ap.add_argument("--serve-check", type=int, default=100,
                    help="requests of a trace whose every result is compared with its solo run (0: skip)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM counter calibration (a 256 MiB device copy, 5 dispatches) for rocprofv3 "
                         "--pmc runs, then exit")
    ap.add_argument("--no-ref", action="store_true", help="skip the full-schedule CPU references (rel-L2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg entirely")
    ap.add_argument("--gen-workers", type=int, default=None,
                    help="processes generating the synthetic inputs (default: up to 16 forked workers; 1 = in-process, "
                         "for runs under a profiler that initialises the GPU before the program starts)")
    args = ap.parse_args()
    args.legs = set(x for x in args.legs.split(",") if x)
    if args.calibrate:
        calibrate()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(slices are independent, so N GPUs would each run their own batch).", file=sys.stderr)
        sys.exit(2)
    workers = args.gen_workers if args.gen_workers is not None else min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    cine, parts, full, split = gen_inputs(args, workers)
    log(f"generated inputs in {time.perf_counter() - t:.1f} s")
    g = gpu_legs(args, cine, parts, full, split)
    c = {} if args.no_cpu else cpu_leg(args, cine[0], full, g)
    it = g["batched"]["iteration"]
    achieved = it["achieved_GBs"]
    tr = load_traffic()
    traffic = tr["hbm_bytes_per_iter"] if tr else None
    line = {
        "metric": METRIC,
        "value": g["batched"]["recons_per_s"],
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["batched"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded phantoms, Gaussian coil maps, variable-density line masks)",
        "config": {"workload": f"NESTA-TV cine slices F={F} C={C} {H}x{W} R={ACCEL:g}, {STAGES} stages x {ITERS} "
                               f"iterations, batch of {args.batch} slices per step, fused batched form in a hipGraph",
                   "batch": args.batch, "frames": F, "coils": C, "matrix": [H, W], "accel": ACCEL,
                   "stages": STAGES, "iters_per_stage": ITERS},
        "latency_ms": {"batched_p50": g["batched"]["p50_ms"], "batched_p99": g["batched"]["p99_ms"],
                       "fused_single_p50": g["fused_single"]["p50_ms"], "fused_single_p99": g["fused_single"]["p99_ms"],
                       "plain_p50": g["plain"]["p50_ms"], "plain_p99": g["plain"]["p99_ms"]},
        "plain_fp32_recons_per_s": g["plain"]["recons_per_s"],
        "fused_single_recons_per_s": g["fused_single"]["recons_per_s"],
        "speedup_vs_plain": g["batched"]["recons_per_s"] / g["plain"]["recons_per_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_nesta_G_t<16> + k_nesta_S2 (one fused NESTA iteration of the whole batch in the "
                               "Gram form: one column pass and its scalar step; HIP events on the library's stream)",
                     "ms_B_plus_S": it["ms_B_plus_S"], "ms_CA": it["ms_CA"],
                     "bytes_alg_per_iter": it["bytes_alg_per_iter"],
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if tr else None},
        "cpu_baseline": c.get("cpu_baseline"),
        "rel_l2": c.get("rel_l2"),
        "device": g["device"],
    }
    if "simple_recon" in g:
        line["simple_recon"] = dict(g["simple_recon"], **c.get("simple_recon", {}))
    if "mc" in g:
        m = dict(g["mc"])
        m["rel_l2_vs_fp64"] = (c.get("rel_l2") or {}).get("motion_compensated")
        m["roofline"] = {"bound": "hbm", "kernel": "k_copy + k_mc_interp (MotionCompensation)",
                         "achieved": m["kernels"]["forward"]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": m["kernels"]["forward"]["frac_hbm"]}
        line["motion_compensated"] = m
    if "registration" in g:
        r = dict(g["registration"])
        cr = c.get("registration")
        if cr:
            r["cpu_baseline"] = cr
            r["speedup_vs_cpu_oracle"] = cr["cpu_s_per_registration_extrapolated"] / (r["ms"] * 1e-3)
        line["registration"] = r
    if "uneven_batch" in g:
        u = dict(g["uneven_batch"])
        u.pop("shapes_FCHW", None)
        u.pop("reports", None)
        if "plain_uneven" in g:
            u["plain_fp32_recons_per_s"] = g["plain_uneven"]["recons_per_s"]
            u["speedup_vs_plain"] = u["recons_per_s"] / g["plain_uneven"]["recons_per_s"]
        line["uneven_batch"] = u
    if "served" in g:
        s = dict(g["served"])
        pm = g.get("plain_served_mix")
        if pm:
            s["plain_fp32_recons_per_s_on_mix"] = pm["recons_per_s"]
            s["speedup_vs_plain"] = s["recons_per_s"] / pm["recons_per_s"]
        line["served"] = s
        if "served_80pct" in g:
            s8 = dict(g["served_80pct"])
            if pm:
                s8["speedup_vs_plain"] = s8["recons_per_s"] / pm["recons_per_s"]
            line["served_80pct_capacity"] = s8
    if "split" in g:
        sp = g["split"]
        line["split_precision"] = {
            "workload": f"{SPLIT_BATCH} slices F={SPLIT_SHAPE['F']} C={SPLIT_SHAPE['C']} "
                        f"{SPLIT_SHAPE['H']}x{SPLIT_SHAPE['W']} R={SPLIT_SHAPE['accel']:g}, {STAGES} stages x {ITERS} "
                        f"iterations, hipGraph",
            "recons_per_s": sp["split"]["recons_per_s"], "ms_per_step": sp["split"]["ms_per_step"],
            "form": "Gram pass (column DFT GEMMs only; line masks)",
            "fp32_fused_recons_per_s": sp["fp32_fused"]["recons_per_s"],
            "iteration": sp["split"]["iteration"], "fp32_fused_iteration": sp["fp32_fused"]["iteration"],
            "rowcol_recons_per_s": sp["split_rowcol"]["recons_per_s"],
            "rowcol_iteration": sp["split_rowcol"]["iteration"],
            "roofline": {"bound": "mfma", "achieved": sp["split"]["iteration"]["mfma_TFLOPs"],
                         "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": sp["split"]["iteration"]["frac_mfma"]},
            "rel_l2": (c.get("rel_l2") or {}).get("split")}
    if "served_burst" in g:
        s = dict(g["served_burst"])
        s["note"] = "same request mix submitted at once: server capacity"
        if g.get("plain_served_mix"):
            s["speedup_vs_plain"] = s["recons_per_s"] / g["plain_served_mix"]["recons_per_s"]
        line["served_burst"] = s
    if "served_burst_large_budget" in g:
        s = dict(g["served_burst_large_budget"])
        s["note"] = "the same burst under a larger HBM budget: capacity when admission is not the tight limit"
        line["served_burst_large_budget"] = s
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
