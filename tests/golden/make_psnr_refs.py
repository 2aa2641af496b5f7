#!/usr/bin/env python3
"""Writes the FP64 oracle's normalised log-density images of the benchmark
workloads (tests/golden/psnr_<cfg>.npz + .json).

For each workload the FP64 chaos game of oracle/oracle.c runs 2^32 binned
samples at the workload's full accumulator size.  Two images are kept, each
as uint16 codes plus a scale: `raw`, refrakt_amd.quality.log_density of the
histogram itself (the best estimate of the converged image the oracle has),
and `de`, the same after the oracle's density estimation (double precision,
the workload's estimator settings), which depends on the absolute sample
count through the estimator radii.

cfg5 (the denoiser's evaluation flames, 2048x2048) keeps the central
window of each frame at full resolution instead (make_cfg5).

cfg4 keeps three frames of the batched animation (make_cfg4).

Usage: python tests/golden/make_psnr_refs.py [cfg1 cfg2 cfg3 cfg4 cfg5] [--threads N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from refrakt_amd import configs, quality  # noqa: E402

SAMPLES = 1 << 32
FUSE = 20
PARTICLES = 1 << 22          # 1024 binned iterations each
SEEDS = {"cfg1": 101, "cfg2": 202, "cfg3": 303}


def make(name: str, threads: int, samples: int = SAMPLES, out_dir: str = HERE) -> None:
    wl = configs.WORKLOADS[name]()
    iters = samples // PARTICLES
    t0 = time.time()
    hist, n = po.chaos_f64(wl.flame, wl.width, wl.height, PARTICLES, FUSE, iters, seed=SEEDS[name],
                           threads=threads)
    t1 = time.time()
    f = wl.flame
    pool = quality.pool_factor(wl.width)
    q_raw, peak_raw = quality.quantize(quality.log_density(hist, n, pool))
    de = po.density(hist.astype(np.float32), f.estimator_radius, f.estimator_min, f.estimator_curve, fp32=False)
    del hist
    t2 = time.time()
    q_de, peak_de = quality.quantize(quality.log_density(de, n, pool))
    np.savez_compressed(os.path.join(out_dir, f"psnr_{name}.npz"), raw=q_raw, de=q_de)
    meta = {"workload": name, "width": wl.width, "height": wl.height, "pool": pool,
            "peak_raw": peak_raw, "peak_de": peak_de,
            "binned": int(n), "attempted": PARTICLES * iters, "particles": PARTICLES, "fuse": FUSE,
            "iters": iters, "seed": SEEDS[name], "threads": threads,
            "estimator": [f.estimator_radius, f.estimator_min, f.estimator_curve],
            "chaos_seconds": round(t1 - t0, 1), "de_seconds": round(t2 - t1, 1),
            "oracle": "orc_chaos_f64 (FP64, xoshiro256** streams) + orc_density (double)"}
    with open(os.path.join(out_dir, f"psnr_{name}.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print(name, json.dumps(meta), flush=True)


def make_cfg5(seed: int, threads: int, samples: int = SAMPLES, out_dir: str = HERE) -> None:
    """cfg5 (denoiser evaluation): the FP64 chaos game of one held-out flame
    at 2048x2048 and 2^32 samples; `raw` is the pool-1 normalised
    log-density image of the central window (configs.CFG5_WINDOW, bin
    layout).  `de_floor_db` in the json is the PSNR of the oracle's own DE
    image against its raw image over the window: the bias that density
    estimation keeps at any sample count."""
    wl = configs.cfg5(seed)
    x0, y0, w, h = configs.CFG5_WINDOW
    iters = samples // PARTICLES
    t0 = time.time()
    hist, n = po.chaos_f64(wl.flame, wl.width, wl.height, PARTICLES, FUSE, iters, seed=seed, threads=threads)
    t1 = time.time()
    f = wl.flame
    npx = wl.width * wl.height
    raw = quality.log_density(hist[y0:y0 + h, x0:x0 + w], n, 1, npx)
    de = po.density(hist.astype(np.float32), f.estimator_radius, f.estimator_min, f.estimator_curve, fp32=False)
    de = quality.image_to_bin_layout(de)
    de_w = quality.log_density(de[y0:y0 + h, x0:x0 + w], n, 1, npx)
    t2 = time.time()
    # a fixed code step of 2^-11 (rms rounding error 1.4e-4, 88 dB below the
    # image peak) keeps the file small; peak_raw = step * 65535 decodes it
    step = 2.0 ** -11
    q_raw = np.round(raw / step).astype(np.uint16)
    q_de = np.round(de_w / step).astype(np.uint16)
    peak_raw = step * 65535.0
    name = wl.name
    np.savez_compressed(os.path.join(out_dir, f"psnr_{name}.npz"), raw=q_raw)
    meta = {"workload": name, "width": wl.width, "height": wl.height, "window": list(configs.CFG5_WINDOW),
            "pool": 1, "peak_raw": peak_raw, "code_step": step, "binned": int(n),
            "attempted": PARTICLES * iters, "particles": PARTICLES, "fuse": FUSE, "iters": iters, "seed": seed,
            "threads": threads, "estimator": [f.estimator_radius, f.estimator_min, f.estimator_curve],
            "de_floor_db": round(quality.psnr(q_de * step, q_raw * step), 2),
            "chaos_seconds": round(t1 - t0, 1), "de_seconds": round(t2 - t1, 1),
            "oracle": "orc_chaos_f64 (FP64, xoshiro256** streams) + orc_density (double)"}
    with open(os.path.join(out_dir, f"psnr_{name}.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print(name, json.dumps(meta), flush=True)


CFG4_FRAMES = (3, 10, 11)     # 512^2 / 12 xforms, 1024^2 / 10 xforms, 1920x1080 / 11 xforms


def make_cfg4(threads: int, out_dir: str = HERE) -> None:
    """cfg4 (batched animation): FP64 chaos game of three of the 64 frames at
    about 1024 samples per pixel; pool 2 (4 for the 1920-wide frame) raw
    log-density images, frame f stored as `f<f>` with its own peak."""
    frames = configs.cfg4()
    arrays, meta = {}, {"workload": "cfg4", "frames": {}}
    for i in CFG4_FRAMES:
        fr = frames[i]
        W, H = fr.width, fr.height
        iters = max(64, (1024 * W * H) // PARTICLES)
        t0 = time.time()
        hist, n = po.chaos_f64(fr.flame, W, H, PARTICLES, FUSE, iters, seed=400 + i, threads=threads)
        pool = 4 if W > 1024 else 2
        q, peak = quality.quantize(quality.log_density(hist, n, pool))
        arrays[f"f{i}"] = q
        meta["frames"][str(i)] = {"width": W, "height": H, "pool": pool, "peak_raw": peak, "binned": int(n),
                                  "attempted": PARTICLES * iters, "xforms": len(fr.flame.xforms),
                                  "gpu_samples": int(fr.samples), "seconds": round(time.time() - t0, 1)}
        print(i, json.dumps(meta["frames"][str(i)]), flush=True)
    meta.update({"particles": PARTICLES, "fuse": FUSE, "threads": threads,
                 "oracle": "orc_chaos_f64 (FP64, xoshiro256** streams)"})
    np.savez_compressed(os.path.join(out_dir, "psnr_cfg4.npz"), **arrays)
    with open(os.path.join(out_dir, "psnr_cfg4.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--samples", type=int, default=SAMPLES)
    ap.add_argument("--out-dir", default=HERE)
    a = ap.parse_args()
    for name in a.configs:
        if name == "cfg4":
            make_cfg4(a.threads, a.out_dir)
        elif name == "cfg5":
            for seed in configs.CFG5_SEEDS:
                make_cfg5(seed, a.threads, a.samples, a.out_dir)
        else:
            make(name, a.threads, a.samples, a.out_dir)


if __name__ == "__main__":
    main()
