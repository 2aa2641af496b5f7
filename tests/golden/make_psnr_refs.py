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

Usage: python tests/golden/make_psnr_refs.py [cfg1 cfg2 cfg3] [--threads N]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--samples", type=int, default=SAMPLES)
    ap.add_argument("--out-dir", default=HERE)
    a = ap.parse_args()
    for name in a.configs:
        make(name, a.threads, a.samples, a.out_dir)


if __name__ == "__main__":
    main()
