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

Every flame is a pure function of its seed (numpy PCG64).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

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


WORKLOADS = {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3}
