#!/usr/bin/env python3
"""Benchmark of the fast engine on MI355X; prints one JSON line.

Metric 1, binned chaos-game iterations/s at 4K (cfg2): 8 xforms + final xform
+ post-affines, 3-5 of 12 variations per xform, 7680x4320 RGBA FP32
accumulator (3840x2160 at supersample 2), 2^24 particles, 20 unbinned
iterations, then 2^32 binned samples.  A step is one draw_to_bins call of 16
passes over all particles (2^28 iterations), so the default 16 timed steps are
the whole 2^32-sample job.  Inputs (flame, particles, histogram) are resident
in HBM before the timed region.

Metric 2, time to 40 dB PSNR against the FP64 oracle (cfg5): held-out flames
of the denoiser's family at 2048x2048, samples per pixel doubling from 1;
plain (samples + density estimation), raw samples and the denoiser scored
against the same oracle image (refrakt_amd/convergence.py).
"""
import argparse
import faulthandler
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = {1: "k_fast<quad>", 2: "k_fast<wave-aggregated>", 3: "k_fast<naive>", 4: "k_fast<LDS bin cache>",
           5: "k_fast<sort staging> + k_sort_rebucket + k_sort_tiles"}
CFG5_SPP = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]
COPY_BYTES = 1 << 29


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
    ap.add_argument("--psnr-only", action="store_true",
                    help="run only the time-to-40-dB measurement (cfg5) and print its JSON (for profiling)")
    ap.add_argument("--no-programs", action="store_true",
                    help="run the precompiled interpreter instead of the flame's compiled program")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy (HBM counter calibration) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc runs that measure roofline.traffic")
    ap.add_argument("--traffic-child", action="store_true",
                    help="(internal) the program profiled for roofline.traffic: 3 copies of 512 MiB, then the "
                         "workload's warmup and two draws")
    return ap.parse_args()


def cpu_baseline(wl, budget_s):
    """The reference's passes (shaders/flame.glsl:41-90: per-workgroup xform,
    per-slot jsf32, shuffle-row gathers, non-atomic float histogram adds)
    restated in FP32 C and spread over the host's cores (orc_ref_passes_mt),
    at the workload's full size: P particles, T temporal samples, a W x H
    float RGBA histogram.  This is the fallback for the reference GL path on
    Mesa llvmpipe, which cannot run on the GPU box: it has no EGL, OSMesa,
    gbm or X server, so no GL context exists there.  Bounded sample: a
    quarter of the workload's particles; timed are the 20 unbinned
    iterations (first run + 19 warmup passes) and as many binned passes as
    fit in about budget_s seconds."""
    from oracle import pyoracle as po
    threads = max(1, min(16, os.cpu_count() or 1))
    W, H, T = wl.width, wl.height, wl.temporal_samples
    # a quarter of the particles bounds the sample (every pass still covers
    # all T temporal samples and the whole histogram)
    P = max(256 * T, wl.particles // 4)
    ppt = P // T
    fc = po.flame_to_c(wl.flame)
    rng = np.random.default_rng(7)
    R = 1024                                     # shuffle rows (src/main.cpp set_sim_parameters)
    shuf = np.empty((R, ppt), np.uint32)
    for r in range(R):
        shuf[r] = rng.permutation(ppt)
    rows = rng.integers(0, R, size=2 * 4096).astype(np.uint32)
    rs = po.jsf_seed_states(P)
    pts = po.hammersley(ppt)
    A = np.zeros((P, 4), np.float32)
    B = np.zeros((P, 4), np.float32)
    bins = np.zeros((H, W, 4), np.float32)
    ss = po.screen_affine(wl.flame, W, H)
    t0 = time.perf_counter()
    po.ref_passes_mt(fc, T, ppt, pts, A, shuf, rs, rows[:2], 1, 1, 1, 1, 0, threads=threads)
    po.ref_passes_mt(fc, T, ppt, A, B, shuf, rs, rows[2:2 + 2 * wl.warmup_passes], wl.warmup_passes, 1, 1, 0, 0,
                     threads=threads)
    t_warm = time.perf_counter() - t0
    cur, nxt = (B, A) if wl.warmup_passes % 2 else (A, B)
    r0 = 2 + 2 * wl.warmup_passes
    t0 = time.perf_counter()
    drawn = po.ref_passes_mt(fc, T, ppt, cur, nxt, shuf, rs, rows[r0:r0 + 2], 1, 1, 0, 0, 1, W, H, ss, bins, threads)
    t_first = time.perf_counter() - t0
    n = int(max(1, min(256, (budget_s - t_warm - t_first) / max(t_first, 1e-3))))
    cur, nxt = nxt, cur
    t0 = time.perf_counter()
    drawn += po.ref_passes_mt(fc, T, ppt, cur, nxt, shuf, rs, rows[r0 + 2:r0 + 2 + 2 * n], n, 1, 0, 0, 1, W, H, ss,
                              bins, threads)
    t_draw = t_first + time.perf_counter() - t0
    passes = n + 1
    return {"value": drawn / t_draw, "unit": "binned samples/s", "cores": threads, "kind": "port",
            "iterations_per_s_incl_fuse": P * (wl.warmup_passes + 1 + passes) / (t_warm + t_draw),
            "sample": f"{wl.name}: FP32 C restatement of the reference pass (shaders/flame.glsl:41-90) on "
                      f"{threads} threads, {P} particles x T={T}, {W}x{H} float RGBA histogram: "
                      f"{wl.warmup_passes + 1} unbinned passes in {t_warm:.1f} s, {passes} binned passes "
                      f"({drawn} binned) in {t_draw:.1f} s; fallback for the reference GL path on llvmpipe "
                      "(no EGL/OSMesa/gbm/X server on the GPU box, so no GL context)"}


def time_to_40db(rf, args, target_db=40.0):
    """Metric 2 on cfg5 for every held-out flame: the curve of
    refrakt_amd.convergence.quality_curve against the FP64 oracle image of
    the frame's central 1024^2 window (tests/golden/psnr_cfg5_<seed>.npz,
    2^32 samples)."""
    import torch
    from refrakt_amd import configs, convergence, denoise, quality
    golden = os.path.join(ROOT, "tests", "golden")
    out = []
    for seed in configs.CFG5_SEEDS:
        wl = configs.cfg5(seed)
        try:
            meta = json.load(open(os.path.join(golden, f"psnr_{wl.name}.json")))
            ref = quality.dequantize(np.load(os.path.join(golden, f"psnr_{wl.name}.npz"))["raw"], meta["peak_raw"])
        except OSError:
            continue
        ref_t = torch.from_numpy(ref).cuda()
        ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=777, engine=rf.RF_ENGINE_FAST,
                         hist_mode=args.hist_mode, block_waves=args.waves, flame_programs=not args.no_programs)
        dn = ctx.denoiser()
        dn_meta = denoise.load_weights(dn)
        res = convergence.quality_curve(ctx, wl, ref_t, meta["window"], CFG5_SPP, dn, target_db)
        res.update({"flame": wl.name, "de_floor_db": meta["de_floor_db"],
                    "reference": f"FP64 oracle, {meta['binned']} binned samples, window {meta['window']}, pool 1"})
        out.append(res)
        dn.close()
        ctx.close()
    if not out:
        return None
    summary = {}
    for k in ("raw", "plain", "denoised"):
        ts = [r["time_to_target"].get(k) for r in out]
        ok = all(t is not None and t.get("bracketed") for t in ts)
        summary[k + "_s"] = float(np.mean([t["wall_s"] for t in ts])) if ok else None
    return {"target_db": target_db, "workload": "cfg5: held-out flames at 2048x2048, 2^22 particles (1 spp per "
            "pass), samples per pixel " + ",".join(map(str, CFG5_SPP)),
            "mean_time_to_target": summary,
            "denoiser": {k: dn_meta.get(k) for k in ("flames", "steps", "spp_in", "zooms", "spp_target",
                                                      "val_gain_db_by_spp")},
            "flames": out}


def _counter_rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        return None
    with open(files[0]) as fh:
        return list(csv.DictReader(fh))


def _short_name(name):
    """Entry-point name of a (possibly demangled, namespaced) kernel name."""
    import re
    m = re.search(r"(rf_prog_w\d+_h\d+|k_\w+|__amd_rocclr_\w+)", name)
    return m.group(1) if m else name


def measure_traffic(args, timeout=180):
    """HBM bytes of one steady-state draw from two rocprofv3 --pmc runs of
    `bench.py --traffic-child` (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the same counter reads for the child's
    512 MiB device copies.  Returns (bytes or None, details)."""
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, {"error": "rocprofv3 not found"}
    tmp = os.environ.get("TMPDIR", "/tmp")
    child = [sys.executable, "-u", os.path.abspath(__file__), "--traffic-child", "--config", args.config,
             "--hist-mode", str(args.hist_mode), "--waves", str(args.waves)]
    if args.no_programs:
        child.append("--no-programs")
    det = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="rf_traffic_", dir=tmp)
        cmd = ["timeout", "-s", "KILL", str(timeout), prof, "--pmc", ctr, "--output-format", "csv", "-d", d,
               "-o", "t", "--"] + child
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, cwd=tmp)
        rows = _counter_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        if not rows:
            return None, {"error": f"{ctr}: no counter output (rc {r.returncode})", "tail": r.stdout[-400:]}
        vals = [(int(x.get("Dispatch_Id", 0)), _short_name(x.get("Kernel_Name", "")),
                 float(x.get("Counter_Value", 0.0))) for x in rows if x.get("Counter_Name", ctr) == ctr]
        vals.sort()
        copies = [v for _, n, v in vals if "copyBuffer" in n and v * 1024 > COPY_BYTES / 8]
        # the last draw: its sort-reset (if any), the last chaos-game launch and
        # the record passes after it
        chaos = [k for k, (_, n, _) in enumerate(vals) if n.startswith(("rf_prog_", "k_fast"))]
        last = {}
        if chaos:
            k0 = chaos[-1]
            if k0 > 0 and vals[k0 - 1][1] == "k_sort_reset":
                last["k_sort_reset"] = vals[k0 - 1][2]
            last[vals[k0][1]] = vals[k0][2]
            for _, n, v in vals[k0 + 1:]:
                if n in ("k_sort_rebucket", "k_sort_tiles"):
                    last[n] = v
        if len(copies) < 3 or not last:
            return None, {"error": f"{ctr}: copies {len(copies)}, draw kernels {sorted(last)}"}
        factor = COPY_BYTES / (np.mean(copies[:3]) * 1024.0)
        det[ctr] = {"factor": round(float(factor), 3), "kib": {k: round(v) for k, v in last.items()},
                    "bytes": float(sum(last.values()) * 1024.0 * factor)}
    return det["FETCH_SIZE"]["bytes"] + det["WRITE_SIZE"]["bytes"], det


def traffic_child(rf, args):
    import torch
    from refrakt_amd import configs
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    wl = configs.WORKLOADS[args.config]()
    ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=12345, engine=rf.RF_ENGINE_FAST,
                     hist_mode=args.hist_mode, block_waves=args.waves, flame_programs=not args.no_programs)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(wl.width, wl.height)
    fh.warmup(wl.warmup_passes)
    for _ in range(3):
        fh.draw_to_bins(hist, wl.draw_passes)
    torch.cuda.synchronize()


def bench_batch(rf, args):
    """cfg4: 64 frames of one flame family (uneven xform counts, sizes, spp) per
    graph replay.  The autotuner picks the launch shape on a 1/16 sample copy
    first; a step is one replay (zero, chaos game, DE + tonemap of all frames)."""
    import torch
    from refrakt_amd import configs
    # the same order as every other path and the tests: torch's GPU state
    # first, then the library's context
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    frames = configs.cfg4()
    ctx = rf.Context(1 << 12, 1, 16, seed=12345, flame_programs=not args.no_programs)
    print("cfg4: creating the batch", file=sys.stderr, flush=True)
    b = ctx.batch([(f.flame, f.width, f.height, f.samples) for f in frames])
    print("cfg4: autotuning", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    cfg, tune_ms = b.tune(1.0 / 16)
    tune_s = time.perf_counter() - t0
    print(f"cfg4: tuned {cfg} in {tune_s:.1f} s; timing", file=sys.stderr, flush=True)
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
    # a crash in native code prints the Python stack of every thread
    faulthandler.enable()
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
    from refrakt_amd import configs, quality
    if args.traffic_child:
        return traffic_child(rf, args)
    if args.config == "cfg4":
        return bench_batch(rf, args)
    if args.psnr_only:
        print(json.dumps(time_to_40db(rf, args)), flush=True)
        return

    wl = configs.WORKLOADS[args.config]()
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=12345, engine=rf.RF_ENGINE_FAST,
                     hist_mode=args.hist_mode, block_waves=args.waves, flame_programs=not args.no_programs)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(wl.width, wl.height)
    fh.warmup(wl.warmup_passes)                # compiles the flame program
    for _ in range(args.warmup):
        fh.draw_to_bins(hist, wl.draw_passes)
    # the unbinned fuse of the job, timed apart (the program is cached now)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fh.warmup(wl.warmup_passes)
    torch.cuda.synchronize()
    warm_s = time.perf_counter() - t0
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
    sort_ovf = fh.sort_overflow() if mode_used == 5 else None
    value = drawn / wall
    attempted = wl.particles * wl.draw_passes * args.steps

    # The log-density image of the timed job against the FP64 oracle's at
    # 2^32 samples (when the 16 default steps are the whole job).
    psnr_fp64 = None
    try:
        meta = json.load(open(os.path.join(ROOT, "tests", "golden", f"psnr_{wl.name}.json")))
        ref = quality.dequantize(np.load(os.path.join(ROOT, "tests", "golden", f"psnr_{wl.name}.npz"))["raw"],
                                 meta["peak_raw"])
        ht = torch.from_numpy(hist.download()).cuda()
        psnr_fp64 = {"db": round(quality.psnr(quality.log_density(ht, drawn, meta["pool"]),
                                              torch.from_numpy(ref).cuda()), 2),
                     "pool": meta["pool"], "oracle_binned": meta["binned"], "binned": drawn}
        del ht
    except OSError:
        pass

    # Atomic ceiling: 16-byte RGBA requests (4 lanes per bin, one sample each)
    # scattered over an L2-resident footprint (2^16 bins = 1 MiB), measured
    # live: the binned-sample rate a per-sample-atomic histogram could reach.
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
    traffic, traffic_det = (None, {"skipped": True})
    if not args.no_traffic:
        traffic, traffic_det = measure_traffic(args)
    roofline = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_counters": traffic_det,
                "kernel": KERNELS.get(mode_used, "k_fast"), "kernel_avg_ms": round(k_avg, 3),
                "note": "achieved = algorithmic bytes (32 B per binned sample: one 16 B RGBA bin read and written; "
                        "+ 24 B per particle) per draw / kernel_avg_ms, the HIP-event time of one draw_to_bins "
                        "call (all its launches); traffic = HBM bytes per steady-state draw measured in this run "
                        "(rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, scaled by the same counters of a 512 MiB copy)"}
    atomic = None
    if probe:
        atomic = {"ceiling": round(probe, 1), "unit": "binned samples/s", "achieved": round(value, 1),
                  "ratio": round(value / probe, 3),
                  "note": "ceiling = one 16-byte memory-side atomic per binned sample (the quad path), measured "
                          "live; the sort path adds tiles in LDS and issues about 1.3 M memory-side atomics per "
                          "2^28-sample draw, so it runs above that ceiling"}
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
    del out
    hist.close()
    fh.close()
    ctx.close()

    job_iters = wl.particles * (wl.warmup_passes + 1 + wl.draw_passes * args.steps)
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
                   "sort_overflow": round(sort_ovf, 6) if sort_ovf is not None else None},
        "iterations_per_s_incl_fuse": job_iters / (warm_s + wall),
        "fuse_ms": round(warm_s * 1e3, 3),
        "roofline": roofline,
        "atomic_ceiling": atomic,
        "psnr_vs_fp64": psnr_fp64,
        "de_tonemap_ms": round(de_ms, 3),
    }
    if not args.no_psnr:
        line["time_to_40db"] = time_to_40db(rf, args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
