"""Learned denoiser: training data from the chaos game itself, training loop,
weights file and inference on a histogram.

The denoiser maps the normalised log-density image of a low-sample histogram,
log10(1 + v * W * H / n) per RGBA channel (bins layout), to that of a
high-sample histogram of the same flame.  Training pairs are rendered by the
batched renderer: random flames of 2-10 xforms mixing 3-5 of the twelve
variations {linear ... bubble} (half of them with post-affines, half with a
final xform) at 512x512 and zoom 1, 2 or 4, 1-32 samples per pixel for the
input (log-uniform; the input carries its noise level in channel 4) and 1024
for the target.  The network, its kernels and Adam live in the library
(rf_denoiser_*); PyTorch only holds the device tensors of the data set.

    python -m refrakt_amd.denoise --flames 2000 --steps 3000 --out weights.npz
    python -m refrakt_amd.denoise --time 4320x7680,1024x1024    (inference timing)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time
from typing import List, Optional

import numpy as np

from .api import RF_HIST_LDS, Context, Flame
from .configs import random_flame

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "denoiser_weights.npz")


def training_flames(n: int, seed: int = 2024, size: int = 512, zooms=(1.0,)) -> List[Flame]:
    """n flames of the family; flame i is shown at zoom zooms[i % len(zooms)]
    (zoom 4 at 512^2 has the pixel scale of the whole view at 2048^2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [random_flame(rng, size, zooms[i % len(zooms)]) for i in range(n)]


def training_spp(n: int, seed: int, lo: float, hi: float) -> np.ndarray:
    """Input sample densities of n training pairs: log-uniform on [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed ^ 0x5bb))
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))


def render_log_density(ctx: Context, flames: List[Flame], W: int, H: int, spp, out, iters: int,
                       chunk: int = 200):
    """Renders every flame at spp samples per pixel (a number, or one per
    flame) with the batched renderer and writes the denoiser input (16-channel
    bf16 log-density and noise level) into out[i] (a torch tensor
    N x H x W x 16, bfloat16, on the GPU)."""
    import torch
    spps = list(spp) if hasattr(spp, "__len__") else [spp] * len(flames)
    for c0 in range(0, len(flames), chunk):
        part = flames[c0:c0 + chunk]
        b = ctx.batch([(f, W, H, max(1, int(s * W * H))) for f, s in zip(part, spps[c0:c0 + chunk])])
        b.set_config(8, RF_HIST_LDS, iters, 20)
        b.run()
        drawn = b.frame_stats()["drawn"]
        for i in range(len(part)):
            n = max(1, int(drawn[i]))
            ctx.log_density_input(b.hist_ptr(i), out[c0 + i].data_ptr(), W * H, W * H / n)
        torch.cuda.synchronize()
        b.close()


def load_weights(dn, path: str = WEIGHTS) -> dict:
    with np.load(path, allow_pickle=False) as z:
        p = z["params"]
        meta = json.loads(str(z["meta"]))
    dn.set_params(p)
    return meta


def save_weights(dn, path: str, meta: dict) -> None:
    np.savez_compressed(path, params=dn.params(), meta=np.array(json.dumps(meta)))


def denoise_histogram(ctx: Context, dn, hist_ptr: int, W: int, H: int, n: int, x_buf=None, y_buf=None):
    """Denoised normalised log-density (H x W x 4 float32 torch tensor, bins
    layout) of a device histogram of n binned samples."""
    import torch
    x = x_buf if x_buf is not None else torch.empty((1, H, W, 16), dtype=torch.bfloat16, device="cuda")
    y = y_buf if y_buf is not None else torch.empty((1, H, W, 4), dtype=torch.float32, device="cuda")
    ctx.log_density_input(hist_ptr, x.data_ptr(), W * H, W * H / max(1, n))
    dn.infer(x.data_ptr(), y.data_ptr(), 1, H, W)
    return y[0]


ZOOMS = (1.0, 2.0, 4.0)


def train(n_flames: int = 2000, steps: int = 3000, batch: int = 8, crop: int = 256, lr: float = 1e-3,
          seed: int = 2024, out: Optional[str] = None, log_every: int = 100, size: int = 512,
          spp_lo: float = 1.0, spp_in_hi: float = 32.0, spp_hi: float = 1024.0, val_flames: int = 64):
    """Pairs: each flame (zoom 1, 2 or 4) rendered at an input density drawn
    log-uniformly from [spp_lo, spp_in_hi] samples per pixel and at spp_hi."""
    import torch
    ctx = Context(1 << 12, 1, 16, seed=seed)
    flames = training_flames(n_flames + val_flames, seed, size, ZOOMS)
    spp_in = training_spp(len(flames), seed, spp_lo, spp_in_hi)
    t0 = time.time()
    X = torch.empty((len(flames), size, size, 16), dtype=torch.bfloat16, device="cuda")
    T = torch.empty_like(X)
    render_log_density(ctx, flames, size, size, spp_in, X, 64)
    render_log_density(ctx, flames, size, size, spp_hi, T, 1024)
    t_data = time.time() - t0
    print(json.dumps({"event": "data", "flames": len(flames), "seconds": round(t_data, 1)}), flush=True)
    dn = ctx.denoiser(seed)
    g = torch.Generator(device="cpu").manual_seed(seed)
    xb = torch.empty((batch, crop, crop, 16), dtype=torch.bfloat16, device="cuda")
    tb = torch.empty_like(xb)
    t0 = time.time()
    hist = []
    for step in range(steps):
        idx = torch.randint(0, n_flames, (batch,), generator=g)
        oy = torch.randint(0, size - crop + 1, (batch,), generator=g)
        ox = torch.randint(0, size - crop + 1, (batch,), generator=g)
        for k in range(batch):
            i, y0, x0 = int(idx[k]), int(oy[k]) // 8 * 8, int(ox[k]) // 8 * 8
            xb[k].copy_(X[i, y0:y0 + crop, x0:x0 + crop])
            tb[k].copy_(T[i, y0:y0 + crop, x0:x0 + crop])
        # cosine decay to 1 % of the initial rate
        a = lr * (0.01 + 0.99 * 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps))))
        want = (step % log_every == 0) or step == steps - 1
        loss = dn.train_step(xb.data_ptr(), tb.data_ptr(), batch, crop, crop, a, want_loss=want)
        if want:
            hist.append((step, loss))
            print(json.dumps({"event": "step", "step": step, "loss": loss, "lr": a,
                              "seconds": round(time.time() - t0, 1)}), flush=True)
    t_train = time.time() - t0
    # validation: held-out flames, full frames, MSE of the input and of the output
    y = torch.empty((1, size, size, 4), dtype=torch.float32, device="cuda")
    mse_in, mse_out, by_spp = [], [], {}
    for i in range(n_flames, len(flames)):
        dn.infer(X[i:i + 1].data_ptr(), y.data_ptr(), 1, size, size)
        t = T[i, :, :, :4].float()
        mse_in.append(float(((X[i, :, :, :4].float() - t) ** 2).mean()))
        mse_out.append(float(((y[0] - t) ** 2).mean()))
        band = "1-4" if spp_in[i] < 4 else ("4-16" if spp_in[i] < 16 else "16-32")
        by_spp.setdefault(band, []).append(10.0 * math.log10(mse_in[-1] / max(mse_out[-1], 1e-12)))
    meta = {"flames": n_flames, "val_flames": val_flames, "size": size, "spp_in": [spp_lo, spp_in_hi],
            "zooms": list(ZOOMS), "spp_target": spp_hi, "input_channels": "log-density RGBA + noise level",
            "steps": steps, "batch": batch, "crop": crop, "lr": lr, "seed": seed,
            "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1),
            "final_loss": hist[-1][1] if hist else None,
            "val_mse_input": float(np.mean(mse_in)), "val_mse_output": float(np.mean(mse_out)),
            "val_gain_db_by_spp": {k: round(float(np.mean(v)), 2) for k, v in sorted(by_spp.items())},
            "n_params": dn.n_params}
    print(json.dumps({"event": "done", **meta}), flush=True)
    if out:
        save_weights(dn, out, meta)
    return dn, meta


def time_inference(sizes, reps: int = 3, seed: int = 1):
    """Host wall time of one inference per (H, W) in `sizes` (after one
    warm-up call), on random bf16 inputs with the committed weights."""
    import torch
    ctx = Context(1 << 12, 1, 16)
    dn = ctx.denoiser(seed)
    load_weights(dn)
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = []
    for H, W in sizes:
        x = (torch.rand((1, H, W, 16), device="cuda", generator=g) * 2).to(torch.bfloat16)
        y = torch.empty((1, H, W, 4), dtype=torch.float32, device="cuda")
        dn.infer(x.data_ptr(), y.data_ptr(), 1, H, W)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            dn.infer(x.data_ptr(), y.data_ptr(), 1, H, W)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        out.append({"event": "infer", "height": H, "width": W, "ms": round(ms, 3)})
        print(json.dumps(out[-1]), flush=True)
    return out


def time_train_step(batch: int = 8, crop: int = 256, reps: int = 10, seed: int = 1):
    """Host wall time of one training step (forward, backward, Adam) on random
    bf16 inputs and targets of batch x crop x crop, after two warm-up steps."""
    import torch
    ctx = Context(1 << 12, 1, 16)
    dn = ctx.denoiser(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = (torch.rand((batch, crop, crop, 16), device="cuda", generator=g) * 2).to(torch.bfloat16)
    t = (torch.rand((batch, crop, crop, 4), device="cuda", generator=g) * 2).to(torch.bfloat16)
    for _ in range(2):
        dn.train_step(x.data_ptr(), t.data_ptr(), batch, crop, crop, 1e-4, want_loss=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dn.train_step(x.data_ptr(), t.data_ptr(), batch, crop, crop, 1e-4, want_loss=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    out = {"event": "train_step", "batch": batch, "crop": crop, "ms": round(ms, 3)}
    print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--time", default=None,
                    help="time inference instead of training, e.g. 4320x7680,1024x1024 (height x width)")
    ap.add_argument("--time-train", default=None,
                    help="time training steps instead of training, e.g. 8x256 (batch x crop)")
    ap.add_argument("--flames", type=int, default=2000)
    ap.add_argument("--val-flames", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.time:
        time_inference([tuple(int(v) for v in s.split("x")) for s in a.time.split(",")])
        return
    if a.time_train:
        b, c = (int(v) for v in a.time_train.split("x"))
        time_train_step(b, c)
        return
    train(a.flames, a.steps, a.batch, a.crop, a.lr, a.seed, a.out, val_flames=a.val_flames)


if __name__ == "__main__":
    main()
