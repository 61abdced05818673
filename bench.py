#!/usr/bin/env python3
"""Benchmark of the gfx950 post-reconstruction filter stage.

One step = the whole in-loop filter chain (plus film grain when the frame has
grain) over one synthetic 10-bit 4:2:0 3840x2160 frame whose input planes are
already resident in HBM.  Prints one JSON line: Mpix/s of displayed samples,
the roofline of the dominant kernel (by kernel time: HIP events the runtime
stamps at each kernel's own start and end) and the CPU baseline.

A second leg times a mixed batch of 64 frames (sizes uniform over 360p..2160p,
8/10-bit 50/50, grain on the 10-bit frames) as 64 sequential per-frame
launches and as one launch per pass (autotuned, hipGraph replay).  A third leg
serves 10,000 jobs from 32 streams with Poisson arrivals (60 fps each) under an
80 GB HBM budget through the continuous-batching server and reports p50/p99
latency; every job's planes are hashed on the device and compared with the
same frame filtered alone.  A second serving run offers the same streams at
5x the frame rate (300 fps each), where jobs arrive while batches run and the
server forms larger batches.

The CPU-baseline leg runs the C restatement under oracle/ on the same seeded
inputs: the 4K frame (one thread, whole frames for ~12 s) and the 64 batch
frames (frame-parallel, one frame per thread).  It is also the checker: the
first 4K oracle frame is compared with the GPU's output and display planes,
and every batch frame's oracle planes with the batched GPU planes
("bit_exact_vs_oracle").

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4k10|1080p10|cif8]
                    [--batch 64] [--batch-steps 5] [--serve-jobs 10000] [--serve-streams 32]
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import rav2d_amd as R  # noqa: E402

HBM_PEAK_GBS = 8000.0
# integer VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD
# every 4 cycles at 2.4 GHz
VALU_PEAK_WINST_S = 1024 * 2.4e9 / 4

CONFIGS = {
    # name: (w, h, bitdepth, sb128, tools)
    "4k10": (3840, 2160, 10, 1, R.TOOL_ALL),
    "1080p10": (1920, 1080, 10, 1, R.TOOL_ALL),
    "cif8": (352, 288, 8, 0, R.TOOL_ALL & ~R.TOOL_GRAIN),
}

PASS_NAMES = ["deblock_v", "deblock_h", "cdef_ccso", "loop_restoration", "film_grain"]
KERNEL_NAMES = ["deblock_v_kernel<u16>", "deblock_h_kernel<u16>", "cdef_kernel<u16>", "lr_kernel<u16>",
                "grain_apply_kernel<u16>"]

# Per-launch counters of the 4k10 workload's kernels from rocprofv3 --pmc runs
# of `bench.py --steps 5 --warmup 2 --no-cpu --batch 0 --serve-jobs 0`
# (profiles/hbm_traffic_4k10_pmc.md): FETCH_SIZE and WRITE_SIZE in KiB per
# dispatch and SQ_INSTS_VALU in wave instructions.  HBM bytes = FETCH x 2.0 +
# WRITE x 1.0 KiB: the factors make the copy probe's counters equal its known
# 256 MiB read and 256 MiB written.  Valid for the kernels profiled there.
PMC_4K10 = {
    "deblock_v": (37574.0, 24420.0, 8.523e6),
    "deblock_h": (26399.5, 24300.0, 8.156e6),
    "cdef_ccso": (14863.0, 24812.8, 19.56e6),
    "loop_restoration": (45991.4, 24300.0, 23.47e6),
    "film_grain": (19949.7, 24301.9, 6.759e6),
}


def pmc_traffic(name):
    f, w, _ = PMC_4K10[name]
    return int(f * 2 * 1024 + w * 1024)


def frame_samples(w, h):
    return w * h + 2 * (w // 2) * (h // 2)


def digest(planes):
    h = hashlib.sha256()
    for a in planes:
        h.update(a.tobytes())
    return h.hexdigest()


def cpu_cores():
    """Threads for the frame-parallel CPU leg: the cores this process may run
    on, at most 16 (the host share that goes with one GPU here)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16)), n


def cpu_baseline(w, h, bd, sb128, tools, seed, gpu_digests, batch_spec, batch_digests, budget_s=12.0):
    """The CPU oracle (scalar restatement of rav2d's chain) as baseline and
    checker.  Single thread: whole frames of the bench workload for ~budget_s;
    the first one is compared with the GPU planes.  Multi-core: the batch
    frames, one frame per thread, each compared with the batched GPU planes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the oracle is only the CPU-baseline leg here
    from concurrent.futures import ThreadPoolExecutor

    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools)
    grain = bool(sf.params.grain_present)

    def chain(src):
        out = O.filter_frame(src)
        disp = O.apply_grain(src, out) if src.params.grain_present else None
        return out, disp

    out, disp = chain(sf)  # warm, and the checked frame
    exact_4k = {"output": digest(out) == gpu_digests[0]}
    if grain:
        exact_4k["display"] = digest(disp) == gpu_digests[1]
    n, t0 = 0, time.perf_counter()
    while True:
        chain(sf)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or n >= 200:
            break
    sf.close()
    res = {"value": round(n * w * h / 1e6 / dt, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
           "sample": "%d whole %dx%d %d-bit frames through the oracle chain%s in %.1f s (one thread)" % (
               n, w, h, bd, " + film grain" if grain else "", dt)}
    exact = {"4k_frame": exact_4k}
    if batch_spec:
        cores, affinity = cpu_cores()

        def one(i):
            c = batch_spec[i]
            src = R.SynthFrame(c[0], c[1], c[2], c[3], c[4], tools=c[5])
            try:
                t = time.perf_counter()
                o, d = chain(src)
                el = time.perf_counter() - t
                ok_o = digest(o) == batch_digests[i][0]
                ok_d = d is None or digest(d) == batch_digests[i][1]
                return ok_o and ok_d, el, ("" if ok_o else "output") + ("" if ok_d else " display")
            finally:
                src.close()

        t1 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            rs = list(ex.map(one, range(len(batch_spec))))
        dtm = time.perf_counter() - t1
        pix = sum(c[0] * c[1] for c in batch_spec)
        exact["batch_frames"] = {"frames": len(rs), "equal": sum(1 for r in rs if r[0]),
                                 "differing": {i: r[2].strip() for i, r in enumerate(rs) if not r[0]}}
        res["multi_core"] = {"value": round(pix / 1e6 / dtm, 3), "unit": "Mpix/s", "cores": cores,
                             "cores_in_affinity": affinity,
                             "sample": "the %d mixed batch frames (%.1f Mpix), one frame per thread, in %.1f s "
                                       "(wall, including each frame's synthetic input generation)" % (
                                           len(rs), pix / 1e6, dtm)}
    return res, exact


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
    n = len(spec)
    pix = sum(c[0] * c[1] for c in spec)
    A = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]  # per-frame form
    B = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]  # batched form
    grain = [bool(sf.params.grain_present) for sf in sfs]

    def per_frame():
        for f in A:
            f.filter()
            f.apply_grain()

    for _ in range(warmup):
        per_frame()
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        per_frame()
    pf_ms = ctx.timer_stop() / steps
    # the per-frame form by frame size: each size's frames launched one by one
    per_size = {}
    for size in BATCH_SIZES:
        idx = [i for i, c in enumerate(spec) if (c[0], c[1]) == size]
        if not idx:
            continue
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(steps):
            for i in idx:
                A[i].filter()
                A[i].apply_grain()
        ms = ctx.timer_stop() / steps
        px = sum(spec[i][0] * spec[i][1] for i in idx)
        per_size["%dx%d" % size] = {"frames": len(idx), "ms": round(ms, 3), "mpix_s": round(px / 1e6 / (ms / 1e3), 1)}

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
    # kernel time of each pass's batched launches (eager run of the same shapes)
    ecfg = b.config()
    ecfg.use_graph = 0
    b.set_config(ecfg)
    ctx.set_profiling(True)
    ctx.pass_times(reset=True)
    for _ in range(steps):
        b.run()
    times = ctx.pass_times(reset=True)
    ctx.set_profiling(False)
    pass_us = {}
    for i, name in enumerate(PASS_NAMES):
        tot, cnt = times[i]
        if cnt:
            pass_us[name] = round(tot / steps * 1e3, 1)  # kernel time per batch run

    dig_a = [(digest(f.download(1)), digest(f.download(2)) if g else None) for f, g in zip(A, grain)]
    dig_b = [(digest(f.download(1)), digest(f.download(2)) if g else None) for f, g in zip(B, grain)]
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
        "per_frame_by_size": per_size,
        "batched_ms": round(bt_ms, 3),
        "per_frame_mpix_s": round(pix / 1e6 / (pf_ms / 1e3), 1),
        "batched_mpix_s": round(pix / 1e6 / (bt_ms / 1e3), 1),
        "speedup": round(pf_ms / bt_ms, 3),
        "batched_bit_exact_vs_per_frame": dig_a == dig_b,
        "autotuned": cfg.as_dict(),
        "autotune_s": round(tune_s, 2),
        "kernel_us_per_batch": pass_us,
    }
    return res, dig_b


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
            # warm-up: every stream's first frame once through the server, so the
            # batched kernels of both bit depths have run before the clock starts
            for i, fs in enumerate(frames):
                srv.submit(fs[0], 1 << 40 | i)
            srv.drain()
            srv.poll(max_results=len(frames))
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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline (and oracle check) leg")
    ap.add_argument("--batch", type=int, default=64, help="frames of the mixed batch leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve-jobs", type=int, default=10000, help="jobs of the serving leg (0 = skip)")
    ap.add_argument("--serve-streams", type=int, default=32)
    ap.add_argument("--serve-fps-high", type=int, default=300, help="frame rate per stream of the high-load serving run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the filter stage has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
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
    gpu_digests = (digest(f.download(1)), digest(f.download(2)) if grain else None)

    # kernel time of every pass over a second run of the same K steps: the
    # runtime hands each pass's kernel a start/stop event pair stamped at the
    # kernel's own start and end (on the stream it runs on)
    ctx.set_profiling(True)
    ctx.pass_times(reset=True)
    for _ in range(args.steps):
        step()
    times = ctx.pass_times(reset=True)
    ctx.set_profiling(False)
    bps = 2 if bd > 8 else 1
    S = frame_samples(w, h)
    alg_bytes = 2 * S * bps  # one read and one write of the frame's planes per pass
    passes = {}
    for i, name in enumerate(PASS_NAMES):
        tot, cnt = times[i]
        if cnt:
            us = tot / cnt * 1e3
            passes[name] = {"kernel": KERNEL_NAMES[i] if bd > 8 else KERNEL_NAMES[i].replace("u16", "u8"),
                            "avg_us": round(us, 2), "launches": cnt,
                            "hbm_frac": round(alg_bytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
            if args.config == "4k10":
                valu = PMC_4K10[name][2]
                passes[name]["valu_frac"] = round(valu / VALU_PEAK_WINST_S / (us * 1e-6), 4)
    # dominant kernel: the one with the longest kernel time
    dom = max(passes, key=lambda k: passes[k]["avg_us"])
    dom_us = passes[dom]["avg_us"]
    achieved = alg_bytes / (dom_us * 1e-6) / 1e9
    frame_bytes = S * bps * (2 + (1 if grain else 0))
    chain_gbs = frame_bytes / (ms_per_step * 1e-3) / 1e9

    batch = serve = None
    batch_spec, batch_digests = None, None
    n_src = max(args.batch, args.serve_streams if args.serve_jobs > 0 else 0)
    if n_src > 0:
        spec = batch_workload(n_src, args.seed)
        print("bench: %d synthetic mixed frames" % n_src, file=sys.stderr, flush=True)
        sfs = make_sources(spec)
        if args.batch > 0:
            print("bench: mixed batch of %d frames" % args.batch, file=sys.stderr, flush=True)
            batch, batch_digests = bench_batch(ctx, spec[:args.batch], sfs[:args.batch], args.batch_steps, 1)
            batch_spec = spec[:args.batch]
        if args.serve_jobs > 0:
            print("bench: serving %d jobs from %d streams" % (args.serve_jobs, args.serve_streams),
                  file=sys.stderr, flush=True)
            serve = bench_serve(spec, sfs, args.serve_streams, 2, 60.0, args.serve_jobs, args.seed, 80 * 10**9)
            # the same streams at 5x the frame rate: batches form while others run
            print("bench: serving at %d fps per stream" % args.serve_fps_high, file=sys.stderr, flush=True)
            serve["high_load"] = bench_serve(spec, sfs, args.serve_streams, 2, float(args.serve_fps_high),
                                             args.serve_jobs, args.seed + 1, 80 * 10**9)
        for x in sfs:
            x.close()

    copy_gbs = ctx.bandwidth_probe(256 << 20, 20)

    cpu = exact = None
    if not args.no_cpu:
        print("bench: CPU baseline and oracle check leg", file=sys.stderr, flush=True)
        cpu, exact = cpu_baseline(w, h, bd, sb128, tools, args.seed, gpu_digests, batch_spec, batch_digests)
        if batch is not None:
            batch["bit_exact_vs_oracle"] = exact["batch_frames"]["equal"] == exact["batch_frames"]["frames"]

    roof = {"bound": "hbm", "kernel": passes[dom]["kernel"], "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(dom) if args.config == "4k10" else None,
            "algorithmic_bytes": alg_bytes, "kernel_us": dom_us,
            "chain_frac": round(chain_gbs / HBM_PEAK_GBS, 4),
            "copy_probe_GBps": round(copy_gbs, 1),
            "frac_of_copy_probe": round(achieved / copy_gbs, 4)}
    if args.config == "4k10":
        roof["valu"] = {"insts_valu": PMC_4K10[dom][2], "peak_winst_per_s": VALU_PEAK_WINST_S,
                        "frac": passes[dom]["valu_frac"]}
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
        "bit_exact_vs_oracle": None if exact is None else all(exact["4k_frame"].values()) and (
            batch is None or batch["bit_exact_vs_oracle"]),
        "oracle_check": exact,
        "config": {"workload": "per-frame launch, %dx%d %d-bit 4:2:0, sb%d, in-loop chain%s" % (
            w, h, bd, 64 << sb128, " + film grain" if grain else ""),
            "wall_ms_timed_region": round(wall, 2),
            "chain_GBps_algorithmic": round(chain_gbs, 1),
            "passes": passes},
        "roofline": roof,
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
