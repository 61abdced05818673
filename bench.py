#!/usr/bin/env python3
"""Benchmark of the gfx950 post-reconstruction filter stage.

One step = the whole in-loop filter chain (plus film grain when the frame has
grain) over one synthetic 10-bit 4:2:0 3840x2160 frame whose input planes are
already resident in HBM.  Prints one JSON line: Mpix/s of displayed samples,
the roofline of the dominant pass (algorithmic bytes = each pass reads and
writes its planes once) and the CPU baseline (the C restatement under oracle/
on a bounded sample of the same frame).

A second leg times a mixed batch of 64 frames (sizes uniform over 360p..2160p,
8/10-bit 50/50, grain on the 10-bit frames) as 64 sequential per-frame
launches and as one launch per pass (autotuned, hipGraph replay), and checks
both give the same planes.  A third leg serves 10,000 jobs from 32 streams
with Poisson arrivals (60 fps each) under an 80 GB HBM budget through the
continuous-batching server and reports p50/p99 latency; every job's planes
are hashed on the device and compared with the same frame filtered alone.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4k10|1080p10|cif8]
                    [--batch 64] [--batch-steps 5] [--serve-jobs 10000] [--serve-streams 32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import rav2d_amd as R  # noqa: E402

HBM_PEAK_GBS = 8000.0

# tools whose GPU passes exist; the CPU baseline runs the same chain
BUILT = R.TOOL_ALL

CONFIGS = {
    # name: (w, h, bitdepth, sb128, tools)
    "4k10": (3840, 2160, 10, 1, R.TOOL_ALL & BUILT),
    "1080p10": (1920, 1080, 10, 1, R.TOOL_ALL & BUILT),
    "cif8": (352, 288, 8, 0, R.TOOL_ALL & ~R.TOOL_GRAIN & BUILT),
}

PASS_NAMES = ["deblock_v", "deblock_h", "cdef_ccso", "loop_restoration", "film_grain"]

# HBM bytes per launch of each pass on the 4k10 workload, from two rocprofv3
# --pmc runs (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0; the factors make the copy
# probe's counters equal its known 256 MiB read / 256 MiB written); see
# profiles/hbm_traffic_4k10_pmc.md.  Valid for the kernels profiled there.
PMC_TRAFFIC_4K10 = {
    "deblock_v": int(37637.2 * 2 * 1024 + 24300.0 * 1024),
    "deblock_h": int(26609.6 * 2 * 1024 + 24300.0 * 1024),
    "cdef_ccso": int(14787.1 * 2 * 1024 + 24801.0 * 1024),
    "loop_restoration": int(46017.7 * 2 * 1024 + 24300.0 * 1024),
    "film_grain": int(19939.8 * 2 * 1024 + 24301.5 * 1024),
}


def frame_samples(w, h):
    return w * h + 2 * (w // 2) * (h // 2)


def cpu_baseline(w, h, bd, sb128, tools, seed, budget_s=12.0):
    """The CPU oracle (scalar restatement of rav2d's chain), one thread, timed on
    a bounded sample of the same workload: whole frames until ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the oracle is only the CPU-baseline leg here

    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools)
    grain = bool(sf.params.grain_present)

    def step():
        out = O.filter_frame(sf)
        if grain:
            O.apply_grain(sf, out)

    step()  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        step()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or n >= 200:
            break
    # frame-parallel: one frame per thread (the oracle runs outside the GIL)
    from concurrent.futures import ThreadPoolExecutor
    cores = max(1, min(os.cpu_count() or 1, 16))
    per = max(1, int(round(budget_s / max(dt / n, 1e-3) * 0.5)))  # frames per thread, ~half the budget

    def worker(_):
        for _ in range(per):
            step()
        return per

    t1 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        nm = sum(ex.map(worker, range(cores)))
    dtm = time.perf_counter() - t1
    sf.close()
    return {"value": round(n * w * h / 1e6 / dt, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": "%d whole %dx%d %d-bit frames through the oracle chain%s in %.1f s" % (
                n, w, h, bd, " + film grain" if grain else "", dt),
            "multi_core": {"value": round(nm * w * h / 1e6 / dtm, 3), "cores": cores,
                           "sample": "%d frames, one per thread, in %.1f s" % (nm, dtm)}}


BATCH_SIZES = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]


def batch_workload(n, seed):
    """Mixed batch: sizes uniform over 360p..2160p, 8/10-bit 50/50, every tool;
    film grain on the 10-bit frames."""
    import random
    rng = random.Random(seed)
    out = []
    for i in range(n):
        w, h = BATCH_SIZES[rng.randrange(len(BATCH_SIZES))]
        bd = 10 if rng.random() < 0.5 else 8
        tools = R.TOOL_ALL if bd == 10 else R.TOOL_ALL & ~R.TOOL_GRAIN
        out.append((w, h, bd, rng.randrange(2), seed * 1000 + i, tools))
    return out


def make_sources(spec):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda c: R.SynthFrame(c[0], c[1], c[2], c[3], c[4], tools=c[5]), spec))


def bench_batch(ctx, spec, sfs, steps, warmup):
    """Mixed batch: per-frame launches vs one launch per pass (graph)."""
    import hashlib
    n = len(spec)
    pix = sum(c[0] * c[1] for c in spec)
    A = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]  # per-frame form
    B = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]  # batched form
    grain = [bool(sf.params.grain_present) for sf in sfs]

    def per_frame():
        for f, g in zip(A, grain):
            f.filter()
            f.apply_grain()

    for _ in range(warmup):
        per_frame()
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        per_frame()
    pf_ms = ctx.timer_stop() / steps

    b = R.GpuBatch(ctx, B)
    t0 = time.perf_counter()
    cfg = b.autotune(iters=3)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        b.run()
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        b.run()
    bt_ms = ctx.timer_stop() / steps
    # per-pass durations of the batched launches (eager run of the same shapes)
    ecfg = b.config()
    ecfg.use_graph = 0
    b.set_config(ecfg)
    ctx.set_profiling(True)
    ctx.pass_times(reset=True)
    for _ in range(steps):
        b.run()
    times = ctx.pass_times(reset=True)
    ctx.set_profiling(False)
    pass_us, pass_bytes = {}, {}
    for i, name in enumerate(PASS_NAMES):
        tot, cnt = times[i]
        if cnt:
            pass_us[name] = round(tot / cnt * 1e3, 1)
            pass_bytes[name] = b.info(i)[1]

    def digest(frames):
        h = hashlib.sha256()
        for f in frames:
            for which in (1, 2):
                for a in f.download(which):
                    h.update(a.tobytes())
        return h.hexdigest()

    exact = digest(A) == digest(B)
    b.close()
    for f in A + B:
        f.close()
    res = {
        "frames": n,
        "mix": {"%dx%d" % s: sum(1 for c in spec if (c[0], c[1]) == s) for s in BATCH_SIZES},
        "ten_bit_frames": sum(1 for c in spec if c[2] == 10),
        "grained_frames": sum(grain),
        "mpix_per_batch": round(pix / 1e6, 2),
        "per_frame_ms": round(pf_ms, 3),
        "batched_ms": round(bt_ms, 3),
        "per_frame_mpix_s": round(pix / 1e6 / (pf_ms / 1e3), 1),
        "batched_mpix_s": round(pix / 1e6 / (bt_ms / 1e3), 1),
        "speedup": round(pf_ms / bt_ms, 3),
        "batched_bit_exact_vs_per_frame": exact,
        "autotuned": cfg.as_dict(),
        "autotune_s": round(tune_s, 2),
        "pass_us": pass_us,
    }
    return res, times, pass_bytes


def bench_serve(spec, sfs, n_streams, slots, fps, n_jobs, seed, budget):
    """Serving: n_streams Poisson streams (stream i = source i of the mixed
    spec), `slots` device-resident frames each, continuous batching, every
    job's planes hashed on the device and checked against the frame alone."""
    import numpy as np
    ctx = R.GpuContext(0, hbm_budget=budget)
    frames, alone = [], []
    try:
        for i in range(n_streams):
            fs = [R.GpuFrame.from_synth(ctx, sfs[i]) for _ in range(slots)]
            frames.append(fs)
            fs[0].filter()
            fs[0].apply_grain()
            alone.append((fs[0].checksum(1), fs[0].checksum(2)))
        held = ctx.bytes_in_use()
        cfg = R.default_serve_config()
        cfg.checksum = 1
        srv = R.GpuServer(ctx, cfg)
        try:
            res, js, jl = srv.simulate(frames, slots, fps, n_jobs, seed)
        finally:
            srv.close()
        lat = np.array([r.t_done - r.t_submit for r in res]) * 1e3
        ok = all(r.status == 0 and (r.checksum[0], r.checksum[1]) == alone[js[j]] for j, r in enumerate(res))
        t0 = min(r.t_submit for r in res)
        t1 = max(r.t_done for r in res)
        pix = sum(spec[js[j]][0] * spec[js[j]][1] for j in range(n_jobs))
        bsz = np.array([r.batch_frames for r in res])
        return {
            "streams": n_streams, "slots_per_stream": slots, "fps_per_stream": fps, "jobs": n_jobs,
            "hbm_budget_GB": budget / 1e9, "hbm_held_GB": round(held / 1e9, 3),
            "max_batch": cfg.max_batch, "max_inflight": cfg.max_inflight,
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3),
            "mpix_s": round(pix / 1e6 / (t1 - t0), 1),
            "offered_mpix_s": round(pix / 1e6 / (max(r.t_submit for r in res) - t0), 1),
            "mean_batch": round(float(bsz.mean()), 2),
            "all_jobs_equal_alone": bool(ok),
        }
    finally:
        for fs in frames:
            for f in fs:
                f.close()
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="4k10", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--batch", type=int, default=64, help="frames of the mixed batch leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve-jobs", type=int, default=10000, help="jobs of the serving leg (0 = skip)")
    ap.add_argument("--serve-streams", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    w, h, bd, sb128, tools = CONFIGS[args.config]
    print("bench: building %s workload %dx%d %d-bit" % (args.config, w, h, bd), file=sys.stderr, flush=True)
    sf = R.SynthFrame(w, h, bd, sb128, args.seed, tools=tools)
    ctx = R.GpuContext(0)
    f = R.GpuFrame.from_synth(ctx, sf)
    grain = bool(sf.params.grain_present)

    def step():
        f.filter()
        if grain:
            f.apply_grain()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    # timed region: K steps, inputs resident in HBM
    wall0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        step()
    ms = ctx.timer_stop()
    wall = (time.perf_counter() - wall0) * 1e3
    ms_per_step = ms / args.steps
    mpix = w * h * args.steps / (ms / 1e3) / 1e6

    # per-pass kernel durations over a second run of the same K steps
    ctx.set_profiling(True)
    ctx.pass_times(reset=True)
    for _ in range(args.steps):
        step()
    times = ctx.pass_times(reset=True)
    ctx.set_profiling(False)
    bps = 2 if bd > 8 else 1
    S = frame_samples(w, h)
    passes = {}
    for i, name in enumerate(PASS_NAMES):
        tot, cnt = times[i]
        if cnt:
            passes[name] = round(tot / cnt * 1e3, 2)  # us per launch
    # dominant kernel: the pass with the largest total time; its algorithmic
    # bytes are one read of its input planes and one write of its output planes
    dom = max(range(5), key=lambda i: times[i][0])
    dom_us = times[dom][0] / max(times[dom][1], 1) * 1e3
    alg_bytes = 2 * S * bps
    achieved = alg_bytes / (dom_us * 1e-6) / 1e9
    frame_bytes = S * bps * (2 + (1 if grain else 0))
    chain_gbs = frame_bytes / (ms_per_step * 1e-3) / 1e9

    batch = serve = None
    n_src = max(args.batch, args.serve_streams if args.serve_jobs > 0 else 0)
    if n_src > 0:
        spec = batch_workload(n_src, args.seed)
        print("bench: %d synthetic mixed frames" % n_src, file=sys.stderr, flush=True)
        sfs = make_sources(spec)
        if args.batch > 0:
            print("bench: mixed batch of %d frames" % args.batch, file=sys.stderr, flush=True)
            batch, _, _ = bench_batch(ctx, spec[:args.batch], sfs[:args.batch], args.batch_steps, 1)
        if args.serve_jobs > 0:
            print("bench: serving %d jobs from %d streams" % (args.serve_jobs, args.serve_streams),
                  file=sys.stderr, flush=True)
            serve = bench_serve(spec, sfs, args.serve_streams, 2, 60.0, args.serve_jobs, args.seed, 80 * 10**9)
        for x in sfs:
            x.close()

    copy_gbs = ctx.bandwidth_probe(256 << 20, 20)

    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline leg", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, h, bd, sb128, tools, args.seed)

    line = {
        "metric": "bit-exact filtered Mpix/s, 10-bit 4K, batched vs per-frame; p99 ms/frame",
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "msamples_per_s": round(frame_samples(w, h) * args.steps / (ms / 1e3) / 1e6, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16" if bd > 8 else "u8",
        "data": "synthetic",
        "config": {"workload": "per-frame launch, %dx%d %d-bit 4:2:0, sb%d, in-loop chain%s" % (
            w, h, bd, 64 << sb128, " + film grain" if grain else ""),
            "passes_built": [n for n in PASS_NAMES if n in passes],
            "wall_ms_timed_region": round(wall, 2),
            "chain_GBps_algorithmic": round(chain_gbs, 1),
            "pass_us": passes},
        "roofline": {"bound": "hbm", "kernel": PASS_NAMES[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": PMC_TRAFFIC_4K10.get(PASS_NAMES[dom]) if args.config == "4k10" else None,
                     "copy_probe_GBps": round(copy_gbs, 1),
                     "frac_of_copy_probe": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "batch": batch,
        "serving": serve,
    }
    f.close()
    ctx.close()
    sf.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
