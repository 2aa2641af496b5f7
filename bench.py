#!/usr/bin/env python3
"""Benchmark: binned chaos-game iterations/s of the fast engine on MI355X.

Workload (default cfg2): 8 xforms + final xform + post-affines, 3-5 of 12
variations per xform, 7680x4320 RGBA FP32 accumulator (3840x2160 at
supersample 2), 2^24 particles, 20 unbinned iterations, then 2^32 binned
samples.  A step is one draw_to_bins call of 16 passes over all particles
(2^28 iterations), so the default 16 timed steps are the whole 2^32-sample
job.  Inputs (flame, particles, histogram) are resident in HBM before the
timed region.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = {1: "k_fast<quad>", 2: "k_fast<wave-aggregated>", 3: "k_fast<naive>", 4: "k_fast<LDS bin cache>",
           5: "k_fast<sort staging> + k_sort_rebucket + k_sort_tiles"}
# HBM bytes per draw launch from rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000
# (scale factors from a 512 MiB copy), profiles/k_fast_cfg2_hbm_counters.md.
# Keyed by (config, hist mode, waves per block); other keys report null.
MEASURED_TRAFFIC = {("cfg2", 0, 8): 5.32e9, ("cfg2", 5, 8): 5.32e9, ("cfg2", 4, 8): 8.76e9, ("cfg2", 1, 4): 9.06e9}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg2", choices=["cfg1", "cfg2", "cfg3", "cfg4"])
    ap.add_argument("--hist-mode", type=int, default=0, help="0 auto, 1 quad, 2 wave-aggregated, 3 naive, 4 LDS bin cache")
    ap.add_argument("--waves", type=int, default=0, help="waves per workgroup: 4, 8, or 0 for the library default (8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-probe", action="store_true")
    ap.add_argument("--no-psnr", action="store_true", help="skip the time-to-40-dB measurement")
    ap.add_argument("--no-programs", action="store_true",
                    help="run the precompiled interpreter instead of the flame's compiled program")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy (HBM counter calibration) and exit")
    return ap.parse_args()


def cpu_baseline(wl, budget_s):
    """The CPU oracle (FP64 chaos game, C threads) on a bounded sample of the workload."""
    from oracle import pyoracle as po
    threads = max(1, min(16, os.cpu_count() or 1))
    W, H = wl.width, wl.height
    # grow the sample until one run takes about budget_s seconds
    fuse, iters = 20, 256
    parts = threads * 256
    while True:
        t0 = time.perf_counter()
        _, n = po.chaos_f64(wl.flame, W // 4, H // 4, parts, fuse, iters, seed=4, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * budget_s or parts >= (1 << 24):
            break
        parts = int(parts * min(16.0, max(2.0, budget_s / max(dt, 1e-3))))
    return {"value": n / dt, "unit": "binned samples/s", "cores": threads, "kind": "port",
            "sample": f"{wl.name} flame, FP64 C oracle (per-particle chaos game), {parts} particles x "
                      f"({fuse} unbinned + {iters} binned) iterations on {threads} threads into a "
                      f"{W // 4}x{H // 4} histogram; {n} binned in {dt:.1f} s"}


def time_to_psnr(rf, wl, target_db, hist_mode, waves, programs=True):
    """Plain method (more samples + DE): wall time from a fresh flame until the
    normalised log-density image of the density-estimated histogram reaches
    target_db against the FP64 oracle's histogram image at 2^32 samples
    (tests/golden/psnr_<cfg>.npz, `raw`).  Samples double from one pass over
    the particles up to the workload's whole job; the clock covers warmup,
    draws and the DE of the image that is scored, not the scoring."""
    import torch
    from refrakt_amd import quality
    golden = os.path.join(ROOT, "tests", "golden")
    try:
        meta = json.load(open(os.path.join(golden, f"psnr_{wl.name}.json")))
        ref = quality.dequantize(np.load(os.path.join(golden, f"psnr_{wl.name}.npz"))["raw"], meta["peak_raw"])
    except OSError:
        return None
    ref_t = torch.from_numpy(ref).cuda()
    ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=777, engine=rf.RF_ENGINE_FAST,
                     hist_mode=hist_mode, block_waves=waves, flame_programs=programs)
    fh = ctx.flame(wl.flame)
    hist_t = torch.zeros((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    hist = ctx.histogram(wl.width, wl.height, device_ptr=hist_t.data_ptr())
    de = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    f = wl.flame
    dep = rf.DEParams(f.estimator_radius, f.estimator_min, f.estimator_curve)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fh.warmup(wl.warmup_passes)
    torch.cuda.synchronize()
    draw_clock = time.perf_counter() - t0      # warmup + every draw so far
    total_passes = wl.draw_passes * wl.draw_calls
    dn = None
    try:
        from refrakt_amd import denoise
        dn = ctx.denoiser()
        dn_meta = denoise.load_weights(dn)
        x_buf = torch.empty((1, wl.height, wl.width, 16), dtype=torch.bfloat16, device="cuda")
        y_buf = torch.empty((1, wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
        # first call sizes the activation buffers; keep that out of the clock
        denoise.denoise_histogram(ctx, dn, hist_t.data_ptr(), wl.width, wl.height, 1, x_buf, y_buf)
    except OSError:
        dn = None
    done, drawn, curve, hit, hit_raw, hit_dn = 0, 0, [], None, None, None
    while done < total_passes:
        n = min(max(done, 1), total_passes - done)   # doubles the samples
        t0 = time.perf_counter()
        drawn += fh.draw_to_bins(hist, n)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        done += n
        draw_clock += t1 - t0
        # Each method's post-processing runs straight after the GPU work it
        # follows (as in a render), before any scoring leaves the GPU idle.
        if dn is not None:
            t3 = time.perf_counter()
            y = denoise.denoise_histogram(ctx, dn, hist_t.data_ptr(), wl.width, wl.height, drawn, x_buf, y_buf)
            torch.cuda.synchronize()
            t4 = time.perf_counter()
        t5 = time.perf_counter()
        ctx.density_estimate(hist, dep, de.data_ptr())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        wall = draw_clock + (t2 - t5)            # one DE of the image that is scored
        p = quality.psnr(quality.log_density(quality.image_to_bin_layout(de), drawn, meta["pool"]), ref_t)
        p_raw = quality.psnr(quality.log_density(hist_t, drawn, meta["pool"]), ref_t)
        curve.append({"samples": drawn, "psnr_db": round(p, 2), "wall_s": round(wall, 4),
                      "psnr_raw_db": round(p_raw, 2), "wall_raw_s": round(draw_clock, 4)})
        if dn is not None:
            dens = torch.clamp(torch.pow(10.0, y) - 1.0, min=0.0)
            p_dn = quality.psnr(quality.log_density(dens, wl.width * wl.height, meta["pool"]), ref_t)
            curve[-1].update({"psnr_denoised_db": round(p_dn, 2), "wall_denoised_s": round(draw_clock + t4 - t3, 4)})
            if hit_dn is None and p_dn >= target_db:
                hit_dn = curve[-1]
        if hit is None and p >= target_db:
            hit = curve[-1]
        if hit_raw is None and p_raw >= target_db:
            hit_raw = curve[-1]
    hist.close()
    if dn is not None:
        dn.close()
    ctx.close()
    return {"target_db": target_db, "plain_s": hit["wall_s"] if hit else None,
            "plain_samples": hit["samples"] if hit else None, "psnr_full_job_db": curve[-1]["psnr_db"],
            "raw_s": hit_raw["wall_raw_s"] if hit_raw else None,
            "raw_samples": hit_raw["samples"] if hit_raw else None,
            "psnr_raw_full_job_db": curve[-1]["psnr_raw_db"],
            "denoised_s": hit_dn["wall_denoised_s"] if hit_dn else None,
            "denoised_samples": hit_dn["samples"] if hit_dn else None,
            "denoiser": ({k: dn_meta[k] for k in ("flames", "steps", "spp_in", "spp_target", "val_mse_input",
                                                   "val_mse_output")} if dn is not None else None),
            "reference": f"FP64 oracle histogram, {meta['binned']} binned samples, pool {meta['pool']}",
            "curve": curve}


def bench_batch(rf, args):
    """cfg4: 64 frames of one flame family (uneven xform counts, sizes, spp) per
    graph replay.  The autotuner picks the launch shape on a 1/16 sample copy
    first; a step is one replay (zero, chaos game, DE + tonemap of all frames)."""
    import torch
    from refrakt_amd import configs
    frames = configs.cfg4()
    ctx = rf.Context(1 << 12, 1, 16, seed=12345, flame_programs=not args.no_programs)
    b = ctx.batch([(f.flame, f.width, f.height, f.samples) for f in frames])
    t0 = time.perf_counter()
    cfg, tune_ms = b.tune(1.0 / 16)
    tune_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = b.frame_stats()
    lat = st["image_done_ms"]
    drawn = int(st["drawn"].sum())
    n = len(frames)
    line = {
        "metric": "batched frames/sec",
        "value": n * args.steps / wall,
        "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded flame family, configs.cfg4)",
        "config": {"workload": f"cfg4: {n} frames, xforms 2-12, sizes 512^2..2048^2, 64-512 spp, one graph replay "
                               "per step", "binned_samples_per_step": drawn,
                   "requested_samples_per_step": int(sum(f.samples for f in frames)),
                   "tuned": cfg, "tune_s": round(tune_s, 2), "parallelism": "single GPU",
                   "flame_program": {"used": b.program_info()[0] == 1, "compile_s": round(b.program_info()[1], 3)}},
        "binned_samples_per_s": drawn * args.steps / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "first": float(lat.min()), "last": float(lat.max()),
                       "note": "device clock, from the start of a replay to each frame's tonemapped image"},
    }
    b.close()
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; --gpus must be 1 (a job would shard by "
                                   "particles, each GPU summing a private histogram, added once at the end)"}))
        sys.exit(2)
    import torch
    if args.copy_calibration:
        src = torch.empty(1 << 27, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 3}))
        return
    import refrakt_amd as rf
    from refrakt_amd import configs
    if args.config == "cfg4":
        return bench_batch(rf, args)

    wl = configs.WORKLOADS[args.config]()
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=12345, engine=rf.RF_ENGINE_FAST,
                     hist_mode=args.hist_mode, block_waves=args.waves, flame_programs=not args.no_programs)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(wl.width, wl.height)
    fh.warmup(wl.warmup_passes)
    for _ in range(args.warmup):
        fh.draw_to_bins(hist, wl.draw_passes)
    hist.zero()
    torch.cuda.synchronize()
    drawn = 0
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        drawn += fh.draw_to_bins(hist, wl.draw_passes)
        kernel_ms.append(ctx.last_draw_ms())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    mode_used, lds_hits = fh.hist_mode()
    prog_used, prog_compile_s = fh.program_info()
    value = drawn / wall
    attempted = wl.particles * wl.draw_passes * args.steps

    # L2 atomic ceiling: 16-byte RGBA requests (4 lanes per bin) scattered over
    # an L2-resident footprint (2^16 bins = 1 MiB), measured live.
    probe = None
    if not args.no_probe:
        samples = 1 << 30
        ms = ctx.atomic_probe(1, 1 << 16, samples)
        probe = samples / (ms / 1e3)
    k_avg = float(np.mean(kernel_ms))
    per_launch_samples = drawn / args.steps
    # Algorithmic HBM bytes of one draw launch: a binned sample reads and
    # writes one 16-byte RGBA bin; each particle's (x, y, color) is read and
    # written once per launch.
    alg_bytes = per_launch_samples * 32 + wl.particles * 24
    hbm_gbs = alg_bytes / (k_avg / 1e3) / 1e9
    atomic_g = per_launch_samples * 4 / (k_avg / 1e3) / 1e9          # G f32 atomic adds/s
    peak_atomic = probe * 4 / 1e9 if probe else None
    roofline = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
                "traffic": MEASURED_TRAFFIC.get((args.config, args.hist_mode, args.waves or 8)),
                "kernel": KERNELS.get(mode_used, "k_fast"), "kernel_avg_ms": round(k_avg, 3),
                "atomic": {"achieved": round(atomic_g, 2), "peak": round(peak_atomic, 2) if peak_atomic else None,
                           "unit": "G f32 atomic adds/s",
                           "frac": round(atomic_g / peak_atomic, 4) if peak_atomic else None},
                "note": "algorithmic bytes = 32 B per binned sample (16 B RGBA read + write) + 24 B per particle; "
                        "kernel_avg_ms = HIP-event time of one draw_to_bins call (all its launches); traffic from "
                        "rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE over the draw's kernels; atomic.achieved counts "
                        "every binned sample's 4 adds against the live probe of 16-byte memory-side atomics "
                        "(above 1: most samples never reach a memory-side atomic)"}
    # DE + tonemap of the finished histogram (HBM-bound pass), timed separately
    out = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    de = rf.DEParams(wl.flame.estimator_radius, wl.flame.estimator_min, wl.flame.estimator_curve)
    tm = rf.TonemapParams(wl.flame.gamma, wl.flame.brightness, wl.flame.vibrancy, 1e-4)
    ctx.density_tonemap(hist, de, tm, out.data_ptr())
    torch.cuda.synchronize()
    td = time.perf_counter()
    ctx.density_tonemap(hist, de, tm, out.data_ptr())
    torch.cuda.synchronize()
    de_ms = (time.perf_counter() - td) * 1e3

    line = {
        "metric": "binned chaos-game iterations/sec at 4K",
        "value": value,
        "unit": "binned samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded flame, configs.py)",
        "config": {"workload": f"{wl.name}: {wl.width}x{wl.height} RGBA f32 accumulator, {wl.particles} particles, "
                               f"{len(wl.flame.xforms)} xforms{' + final' if wl.flame.final_xform else ''}, "
                               f"{wl.draw_passes} passes/step",
                   "binned_samples": drawn, "attempted_samples": attempted, "hist_mode": args.hist_mode,
                   "hist_mode_used": mode_used, "lds_hit_fraction": round(lds_hits, 4) if lds_hits >= 0 else None,
                   "block_waves": args.waves or 8, "parallelism": "single GPU",
                   "flame_program": {"used": bool(prog_used == 1), "compile_s": round(prog_compile_s, 3)},
                   "sort_overflow": round(fh.sort_overflow(), 6) if mode_used == 5 else None},
        "roofline": roofline,
        "de_tonemap_ms": round(de_ms, 3),
    }
    if not args.no_psnr:
        line["time_to_40db"] = time_to_psnr(rf, wl, 40.0, args.hist_mode, args.waves, not args.no_programs)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
