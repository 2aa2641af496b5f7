"""Time to a target image quality: plain (samples + density estimation), raw
(samples only) and denoised (samples + the learned denoiser).

A fresh flame is warmed up and drawn with the samples per pixel doubling
along `spp`; after each step every method's normalised log-density image of
a window of the frame (refrakt_amd.quality.log_density, pool 1, the frame's
normalisation) is scored by PSNR against the same reference image.  The
wall clock of a method at a step is warmup + all draws so far + that
method's post-processing of the step's histogram (the DE kernel, or the
denoiser's input build and inference); scoring is not timed, and neither is
the flame program's compilation and the sizing of the context's buffers (a
first warmup and draw of the same flame structure run before the clock and
are reported).  The plain
method is the reference renderer's (src/main.cpp:490-515 density
estimation after draw_to_bins).
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional, Sequence

from . import quality
from .api import Context, DEParams


def _interp_time(points: List[dict], key: str, target_db: float) -> Optional[dict]:
    """First crossing of target_db along the curve: the bracketing points and
    the wall time interpolated linearly in (log wall, PSNR)."""
    for k, p in enumerate(points):
        if p[key + "_db"] >= target_db:
            if k == 0:
                return {"bracketed": False, "wall_s": p[key + "_s"], "spp": p["spp"]}
            q = points[k - 1]
            a, b = q[key + "_db"], p[key + "_db"]
            f = (target_db - a) / (b - a) if b != a else 1.0
            lt = math.log(q[key + "_s"]) + f * (math.log(p[key + "_s"]) - math.log(q[key + "_s"]))
            return {"bracketed": True, "wall_s": math.exp(lt), "spp_lo": q["spp"], "spp_hi": p["spp"],
                    "db_lo": a, "db_hi": b}
    return None


def _psnr_at(points: List[dict], key: str, wall: float) -> Optional[float]:
    """PSNR of a method at wall time `wall`, interpolated linearly in log wall
    between the curve's points (None outside the curve)."""
    for q, p in zip(points, points[1:]):
        if q[key + "_s"] <= wall <= p[key + "_s"]:
            lq, lp = math.log(q[key + "_s"]), math.log(p[key + "_s"])
            f = (math.log(wall) - lq) / (lp - lq) if lp != lq else 1.0
            return q[key + "_db"] + f * (p[key + "_db"] - q[key + "_db"])
    return None


def quality_curve(ctx: Context, wl, ref, window: Sequence[int], spp: Sequence[int], denoiser=None,
                  target_db: float = 40.0) -> Dict:
    """Runs the protocol above for workload `wl` (one pass = particles / (W H)
    samples per pixel) on context ctx.  ref: the reference log-density image
    of the window (torch tensor on the GPU, h x w x 4); window: (x0, y0, w, h)
    in bins; spp: increasing cumulative samples per pixel; denoiser: an
    rf_denoiser wrapper with weights loaded, or None."""
    import torch
    W, H = wl.width, wl.height
    x0, y0, w, h = window
    npx = W * H
    spp_per_pass = wl.particles / float(npx)
    f = wl.flame
    dep = DEParams(f.estimator_radius, f.estimator_min, f.estimator_curve)
    hist_t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    hist = ctx.histogram(W, H, device_ptr=hist_t.data_ptr())
    de = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    x_buf = y_buf = None
    if denoiser is not None:
        x_buf = torch.empty((1, H, W, 16), dtype=torch.bfloat16, device="cuda")
        y_buf = torch.empty((1, H, W, 4), dtype=torch.float32, device="cuda")
        # the first call sizes the network's activation buffers
        ctx.log_density_input(hist_t.data_ptr(), x_buf.data_ptr(), npx, 1.0)
        denoiser.infer(x_buf.data_ptr(), y_buf.data_ptr(), 1, H, W)
    ctx.density_estimate(hist, dep, de.data_ptr())
    torch.cuda.synchronize()

    # A first flame of this structure compiles its flame program (hiprtc) and,
    # with a two-pass draw, sizes the context's record buffers: reported as
    # cold_warmup_s, kept out of the curve, which starts from a second fresh
    # flame on the warm context (a renderer serving many flames pays these
    # once).
    fh = ctx.flame(f)
    t0 = time.perf_counter()
    fh.warmup(wl.warmup_passes)
    fh.draw_to_bins(hist, 2)
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    hist_t.zero_()
    compile_s = fh.program_info()[1]
    fh.close()
    fh = ctx.flame(f)
    t0 = time.perf_counter()
    fh.warmup(wl.warmup_passes)
    torch.cuda.synchronize()
    clock = time.perf_counter() - t0
    warm = clock
    points = []
    done, drawn = 0, 0
    for s in spp:
        passes = int(round(s / spp_per_pass)) - done
        if passes <= 0:
            continue
        t0 = time.perf_counter()
        drawn += fh.draw_to_bins(hist, passes)
        torch.cuda.synchronize()
        clock += time.perf_counter() - t0
        done += passes
        # each method's post-processing straight after the draw it follows
        t0 = time.perf_counter()
        ctx.density_estimate(hist, dep, de.data_ptr())
        torch.cuda.synchronize()
        de_s = time.perf_counter() - t0
        dn_s = None
        if denoiser is not None:
            t0 = time.perf_counter()
            ctx.log_density_input(hist_t.data_ptr(), x_buf.data_ptr(), npx, npx / max(1, drawn))
            denoiser.infer(x_buf.data_ptr(), y_buf.data_ptr(), 1, H, W)
            torch.cuda.synchronize()
            dn_s = time.perf_counter() - t0
        L_raw = quality.log_density(hist_t[y0:y0 + h, x0:x0 + w], drawn, 1, npx)
        L_de = quality.log_density(quality.image_to_bin_layout(de)[y0:y0 + h, x0:x0 + w], drawn, 1, npx)
        pt = {"spp": round(done * spp_per_pass, 3), "samples": int(drawn),
              "raw_db": round(quality.psnr(L_raw, ref), 2), "raw_s": round(clock, 5),
              "plain_db": round(quality.psnr(L_de, ref), 2), "plain_s": round(clock + de_s, 5)}
        if denoiser is not None:
            L_dn = y_buf[0, y0:y0 + h, x0:x0 + w]
            pt.update({"denoised_db": round(quality.psnr(L_dn, ref), 2), "denoised_s": round(clock + dn_s, 5)})
        points.append(pt)
    hist.close()
    fh.close()
    keys = ["raw", "plain"] + (["denoised"] if denoiser is not None else [])
    out = {"target_db": target_db, "compile_s": round(compile_s, 3), "cold_warmup_s": round(cold, 4),
           "warmup_s": round(warm, 5), "curve": points,
           "time_to_target": {k: _interp_time(points, k, target_db) for k in keys}}
    if denoiser is not None:
        out["matched_wall"] = [{"spp": p["spp"], "wall_s": p["denoised_s"], "denoised_db": p["denoised_db"],
                                "plain_db": _psnr_at(points, "plain", p["denoised_s"]),
                                "raw_db": _psnr_at(points, "raw", p["denoised_s"])}
                               for p in points if p["spp"] <= 16]
    return out
