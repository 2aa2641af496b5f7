"""Learned denoiser: training data from the chaos game itself, training loop,
weights file and inference on a histogram.

The denoiser maps the normalised log-density image of a low-sample histogram,
log10(1 + v * W * H / n) per RGBA channel (bins layout), to that of a
high-sample histogram of the same flame.  Training pairs are rendered by the
batched renderer: random flames of 2-10 xforms mixing 3-5 of the twelve
variations {linear ... bubble} (half of them with post-affines, half with a
final xform) at 512x512, 4 samples per pixel for the input and 1024 for the
target.  The network, its kernels and Adam live in the library
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

from .api import RF_HIST_LDS, Context, Flame, Xform
from .configs import CFG2_VARIATIONS, _rot_scale, smooth_palette

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "denoiser_weights.npz")


def training_flames(n: int, seed: int = 2024, size: int = 512) -> List[Flame]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        xs = []
        post = rng.uniform() < 0.5
        for _ in range(int(rng.integers(2, 11))):
            k = int(rng.integers(3, 6))
            names = list(rng.choice(CFG2_VARIATIONS, size=k, replace=False))
            w = rng.uniform(0, 1, size=k)
            xs.append(Xform(affine=_rot_scale(rng, 0.3, 0.9) + list(rng.uniform(-1, 1, 2)),
                            post=(_rot_scale(rng, 0.8, 1.2) + list(rng.uniform(-0.3, 0.3, 2))) if post else None,
                            variations={a: float(b) for a, b in zip(names, w / w.sum())},
                            weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform(0, 1)), color_speed=0.5,
                            opacity=1.0))
        fin = None
        if rng.uniform() < 0.5:
            fin = Xform(affine=_rot_scale(rng, 0.9, 1.1) + [0.0, 0.0], variations={"linear": 1.0}, weight=0.0,
                        color=float(rng.uniform(0, 1)), color_speed=0.0, opacity=1.0)
        out.append(Flame(xforms=xs, final_xform=fin, palette=smooth_palette(rng), size=[size, size],
                         center=[0.0, 0.0], scale=size / 5.0, rotate=0.0))
    return out


def render_log_density(ctx: Context, flames: List[Flame], W: int, H: int, spp: float, out, iters: int,
                       chunk: int = 200):
    """Renders every flame at spp samples per pixel with the batched renderer
    and writes the denoiser input (16-channel bf16 log-density) into
    out[i] (a torch tensor N x H x W x 16, bfloat16, on the GPU)."""
    import torch
    for c0 in range(0, len(flames), chunk):
        part = flames[c0:c0 + chunk]
        b = ctx.batch([(f, W, H, int(spp * W * H)) for f in part])
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


def train(n_flames: int = 2000, steps: int = 3000, batch: int = 8, crop: int = 256, lr: float = 1e-3,
          seed: int = 2024, out: Optional[str] = None, log_every: int = 100, size: int = 512,
          spp_lo: float = 4.0, spp_hi: float = 1024.0, val_flames: int = 64):
    import torch
    ctx = Context(1 << 12, 1, 16, seed=seed)
    flames = training_flames(n_flames + val_flames, seed, size)
    t0 = time.time()
    X = torch.empty((len(flames), size, size, 16), dtype=torch.bfloat16, device="cuda")
    T = torch.empty_like(X)
    render_log_density(ctx, flames, size, size, spp_lo, X, 64)
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
    mse_in, mse_out = [], []
    for i in range(n_flames, len(flames)):
        dn.infer(X[i:i + 1].data_ptr(), y.data_ptr(), 1, size, size)
        t = T[i, :, :, :4].float()
        mse_in.append(float(((X[i, :, :, :4].float() - t) ** 2).mean()))
        mse_out.append(float(((y[0] - t) ** 2).mean()))
    meta = {"flames": n_flames, "val_flames": val_flames, "size": size, "spp_in": spp_lo, "spp_target": spp_hi,
            "steps": steps, "batch": batch, "crop": crop, "lr": lr, "seed": seed,
            "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1),
            "final_loss": hist[-1][1] if hist else None,
            "val_mse_input": float(np.mean(mse_in)), "val_mse_output": float(np.mean(mse_out)),
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--time", default=None,
                    help="time inference instead of training, e.g. 4320x7680,1024x1024 (height x width)")
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
    train(a.flames, a.steps, a.batch, a.crop, a.lr, a.seed, a.out, val_flames=a.val_flames)


if __name__ == "__main__":
    main()
