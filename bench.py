"""NESTA-TV throughput on one MI355X (recons/s), beside the plain FP32 form and the CPU restatement.

Workload (one "step"): a batch of B independent cine slices, each F=20 frames x C=8 coils x 256x256, variable-density
ky-line mask at acceleration 6 with 16 centre lines, reconstructed with NESTA-TV for 5 continuation stages x 20
inner iterations (stop test off, fixed work).  The batch is advanced by the fused batched engine (two launches per
iteration, replayed from a hipGraph); `value` is whole-job recons/s with inputs resident in HBM.

Also measured in the same run: the plain single-instance FP32 form (the reference's op sequence, host stop test),
the fused single-instance form, per-kernel HIP-event timing of the fused iteration (roofline), the FP32 CPU
restatement on the host cores (cpu_baseline, bounded sample), and the rel-L2 error of the GPU result against the
FP64 CPU oracle and the FP32 CPU restatement on the full schedule.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--batch 16] [--no-ref]
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


def b_alg(F, C, H, W):
    """Algorithmic HBM bytes of one NESTA inner iteration of one slice: read y (8FCN), S (8CN), the u8 mask (FN);
    read x_k, W_k, x_ref and write x_{k+1}, W_{k+1} (5 x 8FN)."""
    N = H * W
    return 8 * N * (F * C + C + 5 * F) + F * N


def _gen(i):
    from opencliper_amd import synth
    sl = synth.make_cine(H=H, W=W, C=C, F=F, accel=ACCEL, slice_id=i)
    return sl.y, sl.S, sl.M


def gen_slices(n, workers):
    with mp.get_context("fork").Pool(max(1, min(workers, n))) as pool:
        return pool.map(_gen, range(n))


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def gpu_leg(args, slices):
    from opencliper_amd import capi
    ctx = capi.Context(0)
    info = ctx.info()
    prm = capi.nesta_params(maxIntIter=STAGES, maxIter=ITERS, stoptest=0)
    out = {"device": info["name"]}
    dev = [(ctx.to_device(y), ctx.to_device(S), ctx.to_device(M), ctx.empty((F, H, W))) for (y, S, M) in slices]
    y0, S0, M0, x0 = dev[0]

    # plain single-instance FP32 form (reference op sequence, host stop test each iteration)
    lat = []
    for r in range(1 + args.plain_steps):
        t = time.perf_counter()
        ctx.nesta(y0, S0, M0, x0, prm, form=capi.OCL_FORM_PLAIN)
        if r:
            lat.append(time.perf_counter() - t)
    out["plain"] = {"recons_per_s": 1.0 / np.mean(lat), "p50_ms": 1e3 * pct(lat, 50), "p99_ms": 1e3 * pct(lat, 99)}
    x_plain = x0.get()

    # fused single-instance form
    lat = []
    for r in range(1 + args.steps):
        t = time.perf_counter()
        ctx.nesta(y0, S0, M0, x0, prm, form=capi.OCL_FORM_FUSED)
        if r:
            lat.append(time.perf_counter() - t)
    out["fused_single"] = {"recons_per_s": 1.0 / np.mean(lat), "p50_ms": 1e3 * pct(lat, 50),
                           "p99_ms": 1e3 * pct(lat, 99)}
    x_fused = x0.get()

    # batched: B slices, one graph of STAGES*ITERS iterations per step
    bt = ctx.batch(dev, prm)
    bt.setup()
    bt.capture(STAGES * ITERS)
    ctx.sync()
    for _ in range(args.warmup):
        bt.setup()
        bt.replay()
        bt.finalize()
        ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        bt.setup()
        bt.replay()
        bt.finalize()
        ctx.sync()
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    B = len(dev)
    out["batched"] = {"recons_per_s": B * args.steps / total, "ms_per_step": 1e3 * total / args.steps,
                      "p50_ms": 1e3 * pct(lat, 50), "p99_ms": 1e3 * pct(lat, 99)}
    x_batch0 = dev[0][3].get()

    # roofline of the fused iteration: HIP events around each launch on the library's stream
    bt.setup()
    ms_b, ms_ca, ms_tot = bt.time_iters(STAGES * ITERS)
    byts = b_alg(F, C, H, W) * B
    it_ms = ms_b + ms_ca
    out["iteration"] = {"ms_B": ms_b, "ms_CA": ms_ca, "ms_iter": it_ms, "span_ms": ms_tot,
                        "bytes_alg_per_iter": byts}
    out["x"] = {"plain": x_plain, "fused": x_fused, "batched0": x_batch0}
    bt.close()
    ctx.close()
    return out


def cpu_leg(args, slice0, gpu_x):
    """Oracle: the FP32 restatement timed on the host cores, and the full-schedule references for slice 0."""
    from oracle import oracle as O
    y, S, M = slice0
    threads = O.num_threads()
    res = {}
    t = time.perf_counter()
    O.nesta(y, S, M, O.params(maxIntIter=1, maxIter=2, stoptest=0), double=False)
    ta = time.perf_counter() - t
    t = time.perf_counter()
    O.nesta(y, S, M, O.params(maxIntIter=1, maxIter=8, stoptest=0), double=False)
    tb = time.perf_counter() - t
    t_iter = (tb - ta) / 6.0
    t_setup = max(0.0, ta - 2 * t_iter)
    t_recon = t_setup + STAGES * ITERS * t_iter
    res["cpu_baseline"] = {"value": 1.0 / t_recon, "unit": "recons/s", "cores": threads, "kind": "port",
                           "sample": f"FP32 restatement (oracle/nesta_oracle.cpp), one config slice, runs of 2 and 8 "
                                     f"inner iterations ({ta + tb:.1f} s of CPU work); setup + 100 iterations "
                                     f"extrapolated linearly ({t_setup:.2f} s setup + {t_iter:.3f} s/iteration)"}
    if not args.no_ref:
        prm = O.params(maxIntIter=STAGES, maxIter=ITERS, stoptest=0)
        x64, _ = O.nesta(y, S, M, prm, double=True)
        x32, _ = O.nesta(y, S, M, prm, double=False)
        res["rel_l2"] = {form: {"vs_fp64": O.rel_l2(x, x64), "vs_fp32_restatement": O.rel_l2(x, x32)}
                         for form, x in gpu_x.items()}
        res["rel_l2"]["fp32_restatement_vs_fp64"] = O.rel_l2(x32, x64)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-ref", action="store_true", help="skip the full-schedule CPU references (rel-L2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg entirely")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(slices are independent, so N GPUs would each run their own batch).", file=sys.stderr)
        sys.exit(2)
    workers = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    slices = gen_slices(args.batch, workers)
    print(f"# generated {args.batch} slices in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    g = gpu_leg(args, slices)
    print(f"# gpu leg done: {json.dumps({k: v for k, v in g.items() if k != 'x'}, default=float)}", file=sys.stderr,
          flush=True)
    c = {} if args.no_cpu else cpu_leg(args, slices[0], g["x"])
    it = g["iteration"]
    achieved = it["bytes_alg_per_iter"] / (it["ms_iter"] * 1e-3) / 1e9
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_nesta_B + k_nesta_CA (one fused NESTA iteration of the whole batch)",
                     "ms_B": it["ms_B"], "ms_CA": it["ms_CA"], "bytes_alg_per_iter": it["bytes_alg_per_iter"]},
        "cpu_baseline": c.get("cpu_baseline"),
        "rel_l2": c.get("rel_l2"),
        "device": g["device"],
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
