"""Seeded synthetic flames for the benchmark workloads.

cfg1  plain method, small: 3 xforms, weights U[0.2,1], affine coefficients
      U[-1,1], each xform linear 0.5 + spherical 0.5, colour U[0,1], a random
      smooth 256-entry palette; 1024x1024 accumulator, 2^20 particles,
      64 iterations x 16 passes, 20 unbinned iterations, DE radius 9 /
      curve 0.4, gamma 2.2.
cfg2  full size: 8 xforms + a final xform, post-affines, 3-5 variations per
      xform from {linear, sinusoidal, spherical, swirl, horseshoe, polar,
      handkerchief, heart, disc, spiral, julia, bubble} with U[0,1] weights
      normalised; 3840x2160 at supersample 2 (7680x4320 accumulator),
      2^24 particles, 2^32 binned samples.
cfg3  atomic contention: 4 strongly contractive xforms (scale U[0.05,0.15],
      offsets within 0.1 of the origin); 8192x8192 accumulator, 2^26
      particles, 2^34 binned samples.
cfg4  batching: 64 animation frames of one seeded flame family; frame i at
      t = i/63 uses the family's first n_i xforms (n_i uneven in 2..12) with
      every affine rotated by 2 pi t times the xform's rotation rate, an
      accumulator of 512^2, 1024^2, 1920x1080 or 2048^2 and 64..512 samples
      per pixel.

Every flame is a pure function of its seed (numpy PCG64).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from .api import Flame, Xform

CFG2_VARIATIONS = ["linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar", "handkerchief", "heart",
                   "disc", "spiral", "julia", "bubble"]


def smooth_palette(rng: np.random.Generator, knots: int = 8) -> np.ndarray:
    """A cyclic palette interpolating `knots` random colours (alpha 1)."""
    ctrl = rng.uniform(0.05, 1.0, size=(knots, 3))
    t = np.arange(256) / 256.0 * knots
    i0 = np.floor(t).astype(int) % knots
    i1 = (i0 + 1) % knots
    f = (t - np.floor(t))[:, None]
    f = f * f * (3 - 2 * f)
    rgb = ctrl[i0] * (1 - f) + ctrl[i1] * f
    pal = np.ones((256, 4), np.float32)
    pal[:, :3] = rgb
    return pal


def _rot_scale(rng, smin, smax):
    th = rng.uniform(0, 2 * math.pi)
    s = rng.uniform(smin, smax)
    c, sn = math.cos(th) * s, math.sin(th) * s
    # flam3 coefs a b c d: x' = a x + c y, y' = b x + d y
    return [c, sn, -sn, c]


def random_flame(rng: np.random.Generator, size: int = 512, zoom: float = 1.0) -> Flame:
    """One flame of the denoiser's family (the cfg2 variation mix): 2-10
    xforms each mixing 3-5 of the twelve cfg2 variations with U[0, 1]
    weights normalised, half of the flames with post-affines, half with a
    final xform, a random smooth palette, square view of side `size` pixels
    showing [-2.5, 2.5]^2 / zoom."""
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
    return Flame(xforms=xs, final_xform=fin, palette=smooth_palette(rng), size=[size, size],
                 center=[0.0, 0.0], scale=size / 5.0 * zoom, rotate=0.0)


@dataclass
class Workload:
    name: str
    flame: Flame
    width: int
    height: int
    particles: int
    temporal_samples: int
    warmup_passes: int      # passes after the first run (first run + warmup = unbinned iterations)
    draw_passes: int        # binned passes per draw call
    draw_calls: int

    @property
    def binned_iterations(self) -> int:
        return self.particles * self.draw_passes * self.draw_calls


def cfg1(seed: int = 1) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = []
    for _ in range(3):
        xs.append(Xform(affine=list(rng.uniform(-1, 1, 6)), variations={"linear": 0.5, "spherical": 0.5},
                        weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform(0, 1)), color_speed=0.5,
                        opacity=1.0))
    f = Flame(xforms=xs, palette=smooth_palette(rng), size=[1024, 1024], center=[0.0, 0.0], scale=1024 / 4.0,
              rotate=0.0, estimator_min=0, estimator_radius=9, estimator_curve=0.4, gamma=2.2, vibrancy=1.0,
              brightness=4.0)
    # 20 unbinned iterations = first run + 19 warmup passes; 16 draws of 64 passes
    return Workload("cfg1", f, 1024, 1024, 1 << 20, 512, 19, 64, 16)


def cfg2(seed: int = 2) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = []
    for _ in range(8):
        k = int(rng.integers(3, 6))
        names = list(rng.choice(CFG2_VARIATIONS, size=k, replace=False))
        w = rng.uniform(0, 1, size=k)
        w = w / w.sum()
        aff = _rot_scale(rng, 0.3, 0.9) + list(rng.uniform(-1, 1, 2))
        post = _rot_scale(rng, 0.8, 1.2) + list(rng.uniform(-0.3, 0.3, 2))
        xs.append(Xform(affine=aff, post=post, variations={n: float(v) for n, v in zip(names, w)},
                        weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform(0, 1)), color_speed=0.5,
                        opacity=1.0))
    fin = Xform(affine=_rot_scale(rng, 0.9, 1.1) + [0.0, 0.0], variations={"linear": 1.0}, weight=0.0,
                color=float(rng.uniform(0, 1)), color_speed=0.0, opacity=1.0)
    f = Flame(xforms=xs, final_xform=fin, palette=smooth_palette(rng), size=[3840, 2160], center=[0.0, 0.0],
              scale=2160 / 5.0, rotate=0.0, estimator_min=0, estimator_radius=9, estimator_curve=0.4, gamma=2.2,
              vibrancy=1.0, brightness=4.0)
    # 2^24 particles x 256 binned passes = 2^32 samples (16 draw calls of 16 passes)
    return Workload("cfg2", f, 7680, 4320, 1 << 24, 512, 19, 16, 16)


def cfg3(seed: int = 3) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = []
    for _ in range(4):
        aff = _rot_scale(rng, 0.05, 0.15) + list(rng.uniform(-0.1, 0.1, 2))
        xs.append(Xform(affine=aff, variations={"linear": 1.0}, weight=float(rng.uniform(0.2, 1.0)),
                        color=float(rng.uniform(0, 1)), color_speed=0.5, opacity=1.0))
    # The attractor lies within ~0.15 of the origin; a view of +-2 units puts it
    # in the central ~4% of the width, so the hits crowd into well under 1% of
    # the pixels.
    f = Flame(xforms=xs, palette=smooth_palette(rng), size=[8192, 8192], center=[0.0, 0.0], scale=8192 / 4.0,
              rotate=0.0, estimator_min=0, estimator_radius=9, estimator_curve=0.4, gamma=2.2, vibrancy=1.0,
              brightness=4.0)
    return Workload("cfg3", f, 8192, 8192, 1 << 26, 512, 19, 16, 16)


@dataclass
class Frame:
    flame: Flame
    width: int
    height: int
    spp: int

    @property
    def samples(self) -> int:
        return self.spp * self.width * self.height


DEGREES_PER_SECOND = 18.0


def rotate_affine(a, deg):
    """flame::rotate_affine (src/flame.hpp:97-109) in float32: the linear part
    [a c; b d] times the rotation [cos -sin; sin cos] on the right,
    a' = a cos + c sin, b' = b cos + d sin, c' = c cos - a sin, d' = d cos - b sin."""
    f = np.float32
    rad = f(0.01745329251) * f(deg)
    sn, cs = f(math.sin(rad)), f(math.cos(rad))
    a = [f(v) for v in a]
    out = [a[0] * cs + a[2] * sn, a[1] * cs + a[3] * sn, a[2] * cs - a[0] * sn, a[3] * cs - a[1] * sn, a[4], a[5]]
    return [float(v) for v in out]


CFG4_SIZES = [(512, 512), (1024, 1024), (1920, 1080), (2048, 2048)]


def cfg4(seed: int = 4, frames: int = 64) -> List[Frame]:
    rng = np.random.Generator(np.random.PCG64(seed))
    fam = []
    for _ in range(12):
        k = int(rng.integers(2, 4))
        names = list(rng.choice(CFG2_VARIATIONS, size=k, replace=False))
        w = rng.uniform(0.2, 1, size=k)
        fam.append({"aff": _rot_scale(rng, 0.35, 0.85) + list(rng.uniform(-0.8, 0.8, 2)),
                    "vars": {n: float(v) for n, v in zip(names, w / w.sum())},
                    "weight": float(rng.uniform(0.2, 1.0)), "color": float(rng.uniform(0, 1)),
                    "rate": float(rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.0))})
    pal = smooth_palette(rng)
    out = []
    for i in range(frames):
        t = i / max(1, frames - 1)
        n = int(rng.integers(2, 13))
        W, H = CFG4_SIZES[int(rng.integers(0, len(CFG4_SIZES)))]
        spp = int(rng.integers(64, 513))
        xs = []
        for x in fam[:n]:
            # the reference's animation clock (src/main.cpp:383-395): 18 degrees
            # per second x rotation_frequency; t in [0, 1] spans 20 s, one turn
            # at rate 1
            aff = rotate_affine(x["aff"], DEGREES_PER_SECOND * 20.0 * t * x["rate"])
            xs.append(Xform(affine=aff, variations=dict(x["vars"]), weight=x["weight"],
                            color=x["color"], color_speed=0.5, opacity=1.0))
        fl = Flame(xforms=xs, palette=pal, size=[W, H], center=[0.0, 0.0], scale=min(W, H) / 4.0, rotate=0.0,
                   estimator_min=0, estimator_radius=9, estimator_curve=0.4, gamma=2.2, vibrancy=1.0,
                   brightness=4.0)
        out.append(Frame(fl, W, H, spp))
    return out


# Held-out flames of the denoiser's family (the training set draws from
# PCG64(2024)), chosen for attractors that fill a good part of the view.
CFG5_SEEDS = (5030, 5070)
# Image-quality window of the 2048^2 frame: the central 1024^2 bins, scored
# at full resolution (pool 1).
CFG5_WINDOW = (512, 512, 1024, 1024)   # x0, y0, width, height in bins


def cfg5(seed: int = CFG5_SEEDS[0]) -> Workload:
    """Denoiser evaluation frame: one held-out flame of the training family
    at 2048x2048, 2^22 particles (one pass = 1 sample per pixel), 20
    unbinned iterations, up to 1024 samples per pixel (16 x 64 passes)."""
    f = random_flame(np.random.Generator(np.random.PCG64(seed)), 2048)
    return Workload(f"cfg5_{seed}", f, 2048, 2048, 1 << 22, 1, 19, 64, 16)


WORKLOADS = {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3, "cfg5": cfg5}
