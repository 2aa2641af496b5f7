"""GPU parity: the HIP path against the CPU oracle, through the C ABI.

Tolerances (stated per test):
  * bin indices, hit counts, drawn counts, particle positions and RNG states of
    affine/linear-only flames: bit-exact;
  * histogram colour sums: relative 1e-5 (float atomics add in any order);
  * single-xform evaluation vs the FP64 oracle: |d| <= 2e-4 * (1 + |ref|) on
    at least 99.5% of points (98% for variations with floor/round/trunc
    discontinuities, where float and double can land on either side);
  * density estimation vs the FP32-restated oracle: 2e-5 of the image maximum;
  * tonemap vs the FP64 oracle: 2e-5 absolute on [0, 1] outputs.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _linear_flame(rf, seed, n=3, final=True, post=True, W=256, H=256):
    rng = np.random.default_rng(seed)
    xs = []
    for _ in range(n):
        th = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(0.35, 0.8)
        aff = [s * math.cos(th), s * math.sin(th), -s * math.sin(th), s * math.cos(th)] + list(rng.uniform(-0.8, 0.8, 2))
        xs.append(rf.Xform(affine=aff, post=list(rng.uniform(-1, 1, 6)) if post and rng.uniform() < 0.5 else None,
                           variations={"linear": float(rng.uniform(0.6, 1.2))}, weight=float(rng.uniform(0.2, 1)),
                           color=float(rng.uniform()), color_speed=0.5, opacity=1.0))
    fin = None
    if final:
        fin = rf.Xform(affine=[0.9, 0.1, -0.1, 0.9, 0.05, -0.02], variations={"linear": 1.0}, color=0.3,
                       color_speed=0.2, opacity=1.0)
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    pal[:, 1] = np.linspace(1, 0, 256)
    pal[:, 2] = 0.5
    return rf.Flame(xforms=xs, final_xform=fin, palette=pal, size=[W, H], center=[0.0, 0.0], scale=W / 5.0)


def _ctx(rf, P, T, engine, hist_mode=1, R=16, seed=7, waves=4):
    return rf.Context(P, T, R, seed=seed, engine=engine, hist_mode=hist_mode, block_waves=waves)


# ------------------------------------------------------------ compat engine
def test_compat_bit_exact_vs_oracle(rf, orc, gpu_available):
    P, T, R, W, H = 1 << 14, 4, 16, 256, 256
    ppt = P // T
    fl = _linear_flame(rf, 11, W=W, H=H)
    ctx = _ctx(rf, P, T, rf.RF_ENGINE_COMPAT, R=R)
    rows_rng = np.random.default_rng(5)
    warm = 3
    wrows = rows_rng.integers(0, R, size=2 * (warm + 1)).astype(np.uint32)
    drows = rows_rng.integers(0, R, size=2 * 5).astype(np.uint32)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(warm, rows=wrows)
    drawn = fh.draw_to_bins(hist, 5, rows=drows)
    g_hist = hist.download()
    g_pos = fh.particles()
    g_rs = ctx.rng_states()

    shuf = ctx.shuffle_rows()
    assert shuf.shape == (R, ppt)
    assert np.array_equal(np.sort(shuf, axis=1), np.tile(np.arange(ppt, dtype=np.uint32), (R, 1)))
    samples = orc.hammersley(ppt)
    assert np.array_equal(ctx.sample_points()[:, :2].view(np.uint32), samples[:, :2].view(np.uint32))

    fc = orc.flame_to_c(fl)
    rs = orc.jsf_seed_states(P)
    bufs = [np.zeros((P, 4), np.float32), np.zeros((P, 4), np.float32)]
    orc.compat_pass(fc, T, ppt, samples, bufs[0], shuf, rs, wrows[0], wrows[1], 1, 1, 1, 0)
    for i in range(warm):
        orc.compat_pass(fc, T, ppt, bufs[i % 2], bufs[(i + 1) % 2], shuf, rs, wrows[2 + 2 * i], wrows[3 + 2 * i],
                        1, 1, 0, 0)
    if warm % 2:
        bufs = [bufs[1], bufs[0]]
    ss = orc.screen_affine(fl, W, H)
    assert np.array_equal(ss, rf.screen_affine(fl, W, H))
    bins = np.zeros((H, W, 4), np.float64)
    hits = np.zeros((H, W), np.uint32)
    o_drawn = 0
    for i in range(5):
        o_drawn += orc.compat_pass(fc, T, ppt, bufs[i % 2], bufs[(i + 1) % 2], shuf, rs, drows[2 * i],
                                   drows[2 * i + 1], 1, 0, 0, 1, W, H, ss, bins, hits)
    final = bufs[1] if 5 % 2 else bufs[0]

    assert drawn == o_drawn > 0
    assert np.array_equal(g_pos.view(np.uint32), final.view(np.uint32)), "particle positions differ"
    assert np.array_equal(g_rs, rs), "jsf32 states differ"
    assert np.array_equal(g_hist[..., 3].astype(np.uint32), hits), "hit counts (bin indices) differ"
    np.testing.assert_allclose(g_hist[..., :3], bins[..., :3], rtol=1e-5, atol=1e-5)


# -------------------------------------------------------------- fast engine
@pytest.mark.parametrize("hist_mode", [1, 2, 3, 4, 5])
def test_fast_bit_exact_vs_oracle_replica(rf, orc, gpu_available, hist_mode):
    P, T, W, H, seed = 1 << 13, 1, 192, 160, 99
    fl = _linear_flame(rf, 21, W=W, H=H)
    ctx = _ctx(rf, P, T, rf.RF_ENGINE_FAST, hist_mode=hist_mode, seed=seed)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(2)
    drawn = fh.draw_to_bins(hist, 6)
    g_hist = hist.download()
    g_pos = fh.particles()
    perms = ctx.fast_perms(4)
    assert np.array_equal(np.sort(perms, axis=1), np.tile(np.arange(256, dtype=np.uint16), (8, 1)))

    fc = orc.flame_to_c(fl)
    ss = orc.screen_affine(fl, W, H)
    pos = np.zeros((P, 4), np.float32)
    orc.fast_run(fc, P // T, P, 4, perms, seed, 0, 3, True, False, W, H, ss, pos)
    bins = np.zeros((H, W, 4), np.float64)
    hits = np.zeros((H, W), np.uint32)
    o_drawn = orc.fast_run(fc, P // T, P, 4, perms, seed, 1, 6, False, True, W, H, ss, pos, bins, hits)
    assert drawn == o_drawn > 0
    assert np.array_equal(g_pos.view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(g_hist[..., 3].astype(np.uint32), hits)
    if hist_mode == 5:
        # the sort path sums tile colours in fixed point, 2^16 - 1 per unit
        # (2^15 - 1 for tiles of more than 65532 records): each sample's
        # colour is rounded by 1 / (2 * 32767) at most
        tol = 1e-5 + 1e-5 * np.abs(bins[..., :3]) + hits[..., None] / 65534.0
        assert (np.abs(g_hist[..., :3].astype(np.float64) - bins[..., :3]) <= tol).all()
    else:
        np.testing.assert_allclose(g_hist[..., :3], bins[..., :3], rtol=1e-5, atol=1e-5)


def test_fast_sort_long_draw_runs_as_subdraws(rf, orc, gpu_available):
    """A sort-mode draw_to_bins call of more than 64 passes runs as chaos
    launches of at most 64 passes (each with the next launch key) that share
    the record buffers: bit-exact against the replica run the same way."""
    P, T, W, H, seed = 1 << 13, 1, 192, 160, 19
    fl = _linear_flame(rf, 29, W=W, H=H)
    ctx = _ctx(rf, P, T, rf.RF_ENGINE_FAST, hist_mode=rf.RF_HIST_SORT, seed=seed)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(2)
    drawn = fh.draw_to_bins(hist, 150)                 # sub-draws of 64, 64 and 22 passes
    g_hist = hist.download()
    g_pos = fh.particles()
    perms = ctx.fast_perms(4)
    fc = orc.flame_to_c(fl)
    ss = orc.screen_affine(fl, W, H)
    pos = np.zeros((P, 4), np.float32)
    orc.fast_run(fc, P // T, P, 4, perms, seed, 0, 3, True, False, W, H, ss, pos)
    hits = np.zeros((H, W), np.uint32)
    o_drawn = 0
    for launch, n in ((1, 64), (2, 64), (3, 22)):
        o_drawn += orc.fast_run(fc, P // T, P, 4, perms, seed, launch, n, False, True, W, H, ss, pos, None, hits)
    assert drawn == o_drawn > 0
    assert np.array_equal(g_pos.view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(g_hist[..., 3].astype(np.uint32), hits)
    # the next call plans from the measured distribution: exact in count, and
    # every record fits its capacity
    d2 = fh.draw_to_bins(hist, 70)
    assert int(hist.download()[..., 3].sum()) == drawn + d2
    assert fh.sort_overflow() == 0.0


# ------------------------------------------------- non-finite particle policy
def _nan_xform(rf, kind):
    """An xform whose output is never finite: a NaN variation weight, or an
    affine that sends every point beyond 1e10 (the badval bound,
    shaders/include/math.glsl)."""
    if kind == "nan":
        return rf.Xform(affine=[0.5, 0.0, 0.0, 0.5, 0.0, 0.0], variations={"linear": float("nan")}, weight=0.35,
                        color=0.9, color_speed=0.5, opacity=1.0)
    return rf.Xform(affine=[1e30, 0.0, 0.0, 1e30, 0.0, 0.0], variations={"linear": 1.0}, weight=0.35, color=0.9,
                    color_speed=0.5, opacity=1.0)


@pytest.mark.parametrize("kind", ["nan", "huge"])
@pytest.mark.parametrize("hist_mode", [1, 5])
def test_fast_nonfinite_particles_restart(rf, orc, gpu_available, kind, hist_mode):
    """The fast engine restarts a particle whose position is NaN or beyond
    1e10 at a uniform point of [-1, 1]^2 and does not draw it that iteration
    (the reference leaves it NaN forever: the reset at shaders/flame.glsl:70 is
    commented out).  Checked bit-exactly against the oracle's replica, which
    applies the same policy, and by its invariants: every particle finite
    afterwards, alpha sum = drawn."""
    P, T, W, H, seed = 1 << 13, 1, 192, 160, 77
    fl = _linear_flame(rf, 23, W=W, H=H)
    fl.xforms.append(_nan_xform(rf, kind))
    print(f"nonfinite restart: kind={kind} hist_mode={hist_mode}", flush=True)
    ctx = _ctx(rf, P, T, rf.RF_ENGINE_FAST, hist_mode=hist_mode, seed=seed)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(2)
    drawn = fh.draw_to_bins(hist, 6)
    g_hist = hist.download()
    g_pos = fh.particles()
    perms = ctx.fast_perms(4)
    assert np.isfinite(g_pos).all() and (np.abs(g_pos[:, :2]) <= 1e10).all()
    assert int(g_hist[..., 3].sum(dtype=np.float64)) == drawn > 0

    fc = orc.flame_to_c(fl)
    ss = orc.screen_affine(fl, W, H)
    pos = np.zeros((P, 4), np.float32)
    orc.fast_run(fc, P // T, P, 4, perms, seed, 0, 3, True, False, W, H, ss, pos)
    hits = np.zeros((H, W), np.uint32)
    o_drawn = orc.fast_run(fc, P // T, P, 4, perms, seed, 1, 6, False, True, W, H, ss, pos, None, hits)
    assert drawn == o_drawn
    assert np.array_equal(g_pos.view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(g_hist[..., 3].astype(np.uint32), hits)


def test_fast_only_nonfinite_xform_draws_nothing(rf, gpu_available):
    """A flame whose one xform never yields a finite point: every iteration
    restarts every particle, nothing is drawn, and the particles end inside
    [-1, 1]^2."""
    W, H = 128, 128
    fl = _linear_flame(rf, 25, n=1, final=False, post=False, W=W, H=H)
    fl.xforms[0] = _nan_xform(rf, "nan")
    ctx = _ctx(rf, 1 << 13, 1, rf.RF_ENGINE_FAST, hist_mode=1, seed=3)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(3)
    assert fh.draw_to_bins(hist, 4) == 0
    assert not hist.download().any()
    p = fh.particles()
    assert np.isfinite(p).all() and (np.abs(p[:, :2]) <= 1.0).all()


def test_auto_probe_runs_again_after_rewarmup(rf, gpu_available):
    """RF_HIST_AUTO measures the LDS table on the first pass of a flame's
    first draw after every warmup (an edited flame may spread differently)."""
    W, H = 640, 480
    fl = _linear_flame(rf, 27, W=W, H=H)
    ctx = _ctx(rf, 1 << 16, 1, rf.RF_ENGINE_FAST, hist_mode=0, seed=5, waves=8)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    for k in range(2):
        fh.warmup(3)
        mode, frac = fh.hist_mode()
        assert mode == 4 and frac < 0, (k, mode, frac)      # LDS table, nothing measured yet
        assert fh.draw_to_bins(hist, 4) > 0
        mode, frac = fh.hist_mode()
        assert frac >= 0 and mode == 5, (k, mode, frac)     # probed, moved to the sort path


# --------------------------------------------------------------- variations
DISCONTINUOUS = {"rings", "fan", "fan2", "rings2", "julian", "juliascope", "pie", "ngon", "rectangles", "boarders",
                 "cell", "cpow", "secant2", "twintrian", "bipolar", "oscope", "julia", "loonie", "bent", "bent2",
                 "tangent", "popcorn", "rays", "arch", "super_shape", "flower", "conic", "disc2", "curl", "foci",
                 "coth", "mobius", "perspective", "cross", "edisc"}


def _var_params(name, rng):
    pr = {
        "blob_high": 1.2, "blob_low": 0.4, "blob_waves": 3.0, "pdj_a": 1.1, "pdj_b": -0.7, "pdj_c": 0.9,
        "pdj_d": 1.6, "fan2_x": 0.6, "fan2_y": 0.3, "rings2_val": 0.7, "perspective_angle": 0.4,
        "perspective_dist": 2.5, "julian_power": 5.0, "julian_dist": 1.3, "juliascope_power": 3.0,
        "juliascope_dist": 0.8, "radial_blur_angle": 0.6, "pie_slices": 6.0, "pie_rotation": 0.5,
        "pie_thickness": 0.4, "ngon_power": 2.0, "ngon_sides": 5.0, "ngon_corners": 1.0, "ngon_circle": 0.5,
        "curl_c1": 0.4, "curl_c2": 0.2, "rectangles_x": 0.6, "rectangles_y": 0.4, "disc2_rot": 0.7,
        "disc2_twist": 1.1, "super_shape_n1": 2.0, "super_shape_n2": 1.5, "super_shape_n3": 1.2,
        "super_shape_m": 4.0, "super_shape_holes": 0.2, "super_shape_rnd": 0.3, "flower_holes": 0.3,
        "flower_petals": 5.0, "conic_holes": 0.2, "conic_eccentricity": 0.7, "parabola_height": 0.8,
        "parabola_width": 1.2, "bent2_x": 1.4, "bent2_y": 0.6, "bipolar_shift": 0.3, "cell_size": 0.7,
        "cpow_power": 3.0, "cpow_r": 1.2, "cpow_i": 0.3, "curve_xlength": 1.1, "curve_ylength": 0.8,
        "curve_xamp": 0.5, "curve_yamp": -0.4, "oscope_separation": 0.5, "oscope_frequency": 1.5,
        "oscope_amplitude": 0.8, "oscope_damping": 0.3, "waves2_scalex": 0.4, "waves2_scaley": 0.3,
        "waves2_freqx": 2.0, "waves2_freqy": 1.5, "auger_freq": 3.0, "auger_scale": 0.5, "auger_sym": 0.4,
        "auger_weight": 0.6, "flux_spread": 0.5, "mobius_re_a": 0.9, "mobius_re_b": 0.2, "mobius_re_c": 0.3,
        "mobius_re_d": 1.1, "mobius_im_a": 0.1, "mobius_im_b": -0.3, "mobius_im_c": 0.25, "mobius_im_d": 0.05,
    }
    return pr


def test_every_variation_vs_fp64_oracle(rf, orc, gpu_available):
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    rng = np.random.default_rng(3)
    n = 4096
    failures = []
    names = rf.variation_names()
    assert sorted(names) == sorted(orc.lib().orc_var_name(i).decode() for i in range(orc.lib().orc_var_count()))
    for name in names:
        prm = {p: v for p, v in _var_params(name, rng).items() if p in rf.variation_params(name)}
        xf = rf.Xform(affine=[0.8, 0.3, -0.2, 0.9, 0.1, -0.05], post=[1.0, 0.1, 0.0, 0.9, 0.02, 0.0],
                      variations={name: 0.7}, var_param=prm, color=0.4, color_speed=0.3, opacity=0.8)
        if name == "pre_blur":  # a pre_xform needs a regular term to produce a point
            xf.variations["linear"] = 1.0
        pts = np.concatenate([rng.uniform(-1.5, 1.5, (n, 2)), rng.uniform(0, 1, (n, 1))], axis=1).astype(np.float32)
        rands = rng.uniform(0, 1, (n, 16)).astype(np.float32)
        g = ctx.eval_xform(xf, pts, rands)
        o = orc.apply_f64(xf, pts.astype(np.float64), rands)
        fin = np.isfinite(o[:, :2]).all(axis=1) & (np.abs(o[:, :2]) < 1e4).all(axis=1)
        err = np.abs(g[:, :2].astype(np.float64) - o[:, :2]) / (1.0 + np.abs(o[:, :2]))
        ok = (err <= 2e-4).all(axis=1) & fin
        frac = ok.sum() / max(fin.sum(), 1)
        need = 0.98 if name in DISCONTINUOUS else 0.995
        if frac < need:
            failures.append((name, float(frac)))
        np.testing.assert_allclose(g[:, 2], o[:, 2], rtol=1e-6, atol=1e-6)
        assert np.all(g[:, 3] == np.float32(0.8))
    assert not failures, failures


# Error budget of the fast engine's sine (the argument in revolutions, x / 2 pi
# rounded to float, into the hardware v_sin_f32): the reduction's rounding
# grows with |x|, so the stated bound per point is 1e-6 + 1.5e-7 |x| (GLSL
# specifies sin only on [-pi, pi], to 2^-11).  Measured maxima: 4.4e-7 on
# [-pi, pi], 7.8e-7 on [-10, 10], 6.8e-6 on [-100, 100], 8.7e-5 on
# [-1000, 1000].
@pytest.mark.parametrize("xmax", [np.pi, 10.0, 100.0, 1000.0])
def test_sine_error_budget_per_range(rf, gpu_available, xmax):
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    rng = np.random.default_rng(int(xmax))
    n = 65536
    pts = np.zeros((n, 3), np.float32)
    pts[:, 0] = rng.uniform(-xmax, xmax, n).astype(np.float32)
    pts[:, 1] = rng.uniform(-xmax, xmax, n).astype(np.float32)
    rands = np.zeros((n, 16), np.float32)
    xf = rf.Xform(affine=[1.0, 0.0, 0.0, 1.0, 0.0, 0.0], variations={"sinusoidal": 1.0}, color=0.0,
                  color_speed=0.0, opacity=1.0)
    g = ctx.eval_xform(xf, pts, rands)
    x = pts[:, :2].astype(np.float64)
    err = np.abs(g[:, :2].astype(np.float64) - np.sin(x))
    bound = 1e-6 + 1.5e-7 * np.abs(x)
    assert (err <= bound).all(), (xmax, float(err.max()), float((err / bound).max()))


# ------------------------------------------------------------- DE, tonemap
def _sparse_hist(rng, W, H, density=0.3):
    h = np.zeros((H, W, 4), np.float32)
    m = rng.uniform(size=(H, W)) < density
    a = np.exp(rng.uniform(-1, 7, size=(H, W))).astype(np.float32)
    h[..., 3] = np.where(m, a, 0)
    h[..., :3] = (h[..., 3:4] * rng.uniform(0.1, 1.0, size=(H, W, 3))).astype(np.float32)
    return h


# R <= 11: 32 x 32 tiles; 12-15: 32 x 8 tiles; beyond: the global gather;
# rmin > R: the splat span is rmin
@pytest.mark.parametrize("R,rmin,curve", [(9, 0, 0.4), (3, 1, 0.6), (0, 0, 0.4), (30, 0, 0.5), (11, 0, 0.3),
                                          (13, 0, 0.45), (2, 5, 0.4)])
def test_density_estimation_vs_oracle(rf, orc, gpu_available, R, rmin, curve):
    import torch
    rng = np.random.default_rng(R + 17)
    W, H = 97, 75
    h = _sparse_hist(rng, W, H)
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    hist = ctx.histogram(W, H)
    hist.upload(h)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    ctx.density_estimate(hist, rf.DEParams(R, rmin, curve), out.data_ptr())
    torch.cuda.synchronize()
    g = out.cpu().numpy()
    ref = orc.density(h, R, rmin, curve, fp32=True)
    tol = 2e-5 * np.abs(ref).max()
    assert np.abs(g - ref).max() <= tol


def test_density_identity_layout(rf, orc, gpu_available):
    """R = 0: each bin lands on texel (px - 1, H - 1 - py); column 0 is clipped."""
    import torch
    W, H = 40, 30
    h = np.zeros((H, W, 4), np.float32)
    h[3, 7] = [1, 2, 3, 4]
    h[10, 0] = [5, 5, 5, 5]
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    hist = ctx.histogram(W, H)
    hist.upload(h)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    ctx.density_estimate(hist, rf.DEParams(0, 0, 0.4), out.data_ptr())
    g = out.cpu().numpy()
    exp = np.zeros_like(g)
    exp[H - 1 - 3, 7 - 1] = [1, 2, 3, 4]
    assert np.array_equal(g, exp)


def test_tonemap_vs_fp64_oracle(rf, orc, gpu_available):
    import torch
    rng = np.random.default_rng(8)
    W, H = 64, 48
    img = np.abs(rng.normal(size=(H, W, 4))).astype(np.float32) * 50
    img[0, :5] = 0
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    for gamma, bright, vib in [(2.2, 4.0, 1.0), (4.0, 29.7, 0.0), (1.5, 10.0, 0.5)]:
        src = torch.from_numpy(img).cuda()
        dst = torch.zeros_like(src)
        ctx.tonemap(src.data_ptr(), dst.data_ptr(), W, H, rf.TonemapParams(gamma, bright, vib, 1e-4))
        g = dst.cpu().numpy()
        ref = orc.tonemap(img.astype(np.float64), gamma, bright, vib, 1e-4)
        assert np.abs(g - ref).max() <= 2e-5


def test_fused_density_tonemap_matches_split(rf, gpu_available):
    import torch
    rng = np.random.default_rng(2)
    W, H = 80, 64
    h = _sparse_hist(rng, W, H)
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    hist = ctx.histogram(W, H)
    hist.upload(h)
    de = rf.DEParams(9, 0, 0.4)
    tm = rf.TonemapParams(2.2, 4.0, 1.0, 1e-4)
    a = torch.zeros((H, W, 4), device="cuda")
    b = torch.zeros((H, W, 4), device="cuda")
    fused = torch.zeros((H, W, 4), device="cuda")
    ctx.density_estimate(hist, de, a.data_ptr())
    ctx.tonemap(a.data_ptr(), b.data_ptr(), W, H, tm)
    ctx.density_tonemap(hist, de, tm, fused.data_ptr())
    assert torch.equal(b, fused)


# -------------------------------------------------------- conservation, size
@pytest.mark.parametrize("hist_mode", [1, 2, 4, 0])
def test_contention_counts_conserved(rf, gpu_available, hist_mode):
    """A strongly contractive flame: every binned sample adds exactly 1 to alpha."""
    from refrakt_amd import configs
    wl = configs.cfg3()
    W = H = 1024
    wl.flame.size = [W, H]
    wl.flame.scale = W / 4.0
    ctx = _ctx(rf, 1 << 18, 1, rf.RF_ENGINE_FAST, hist_mode=hist_mode)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(W, H)
    fh.warmup(19)
    drawn = sum(fh.draw_to_bins(hist, 8) for _ in range(4))
    g = hist.download()
    assert g[..., 3].max() < 2 ** 24
    assert drawn == (1 << 18) * 32
    assert int(g[..., 3].astype(np.float64).sum()) == drawn
    if hist_mode in (0, 4):
        mode, frac = fh.hist_mode()
        assert mode == 4 and frac > 0.95  # a few dozen hot bins: nearly every sample re-hits its LDS slot


@pytest.mark.parametrize("waves", [4, 8])
def test_lds_cache_matches_quad_over_many_batches(rf, gpu_available, waves):
    """2^20 particles: each workgroup of the LDS mode runs several batches and
    flushes once; trajectories and alpha counts equal the quad mode's exactly."""
    from refrakt_amd import configs
    wl = configs.cfg1()
    out = {}
    for mode in (1, 4):
        ctx = _ctx(rf, 1 << 20, 1, rf.RF_ENGINE_FAST, hist_mode=mode, seed=7, waves=waves)
        fh = ctx.flame(wl.flame)
        hist = ctx.histogram(wl.width, wl.height)
        fh.warmup(19)
        drawn = fh.draw_to_bins(hist, 8)
        out[mode] = (drawn, hist.download(), fh.particles(), fh.hist_mode())
    d1, h1, p1, _ = out[1]
    d4, h4, p4, (m4, frac) = out[4]
    assert d1 == d4 > 0
    assert np.array_equal(p1.view(np.uint32), p4.view(np.uint32))
    assert np.array_equal(h1[..., 3], h4[..., 3])
    # float32 sums of up to ~3e4 terms in different orders: up to n * 2^-24 apart
    np.testing.assert_allclose(h4[..., :3], h1[..., :3], rtol=2e-3, atol=1e-3)
    assert m4 == 4 and 0.2 < frac <= 1.0


def _square_flame(rf, W, opacity=1.0):
    xs = [rf.Xform(affine=[0.5, 0.0, 0.0, 0.5, ex, fy], variations={"linear": 1.0}, weight=1.0, color=0.25 * k,
                   color_speed=0.5, opacity=opacity)
          for k, (ex, fy) in enumerate([(-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5)])]
    return rf.Flame(xforms=xs, palette=np.ones((256, 4), np.float32), size=[W, W], center=[0.0, 0.0],
                    scale=W / 2.0, rotate=0.0)


@pytest.mark.parametrize("opacity,expect", [(1.0, 5), (0.75, 1)])
def test_auto_mode_leaves_lds_on_a_spread_footprint(rf, gpu_available, opacity, expect):
    """Four half-scale maps tile the square, so hits spread evenly over all
    2048^2 bins and the LDS tables catch almost none: AUTO switches to the
    sort path after one draw (or to quad when weights are not all 1) and the
    histogram stays exact across the switch."""
    W = H = 2048
    ctx = _ctx(rf, 1 << 20, 1, rf.RF_ENGINE_FAST, hist_mode=0, waves=8)
    fh = ctx.flame(_square_flame(rf, W, opacity))
    hist = ctx.histogram(W, H)
    fh.warmup(19)
    assert fh.hist_mode()[0] == 4
    d = fh.draw_to_bins(hist, 4)
    mode, frac = fh.hist_mode()
    assert 0.0 <= frac < 0.02 and mode == expect
    d += fh.draw_to_bins(hist, 4)
    assert fh.hist_mode()[0] == expect
    g = hist.download()[..., 3]
    assert d == (1 << 20) * 8
    assert abs(float(g.astype(np.float64).sum()) - d * opacity) <= 1e-6 * d
    assert (g > 0).mean() > 0.8  # 2 hits per bin on average: 1 - e^-2 = 86 % of bins are reached
    if expect == 5:
        assert 0.0 <= fh.sort_overflow() < 0.01


@pytest.mark.parametrize("cfg,size", [("cfg2", (1920, 1080)), ("cfg1", (1024, 1024)), ("cfg3", (2048, 2048))])
def test_sort_path_matches_quad(rf, gpu_available, cfg, size):
    """The sort path adds exactly the samples the quad path adds (same
    trajectories: alpha counts equal), over two draws (the second sized from
    the first), including flames whose hits overflow the first draw's
    uniform capacities (cfg1, cfg3: everything lands in a few tiles)."""
    from refrakt_amd import configs
    wl = configs.WORKLOADS[cfg]()
    W, H = size
    wl.flame.size = [W, H]
    wl.flame.scale = wl.flame.scale * W / wl.width
    out = {}
    for mode in (1, 5):
        ctx = _ctx(rf, 1 << 20, 1, rf.RF_ENGINE_FAST, hist_mode=mode, seed=11, waves=8)
        fh = ctx.flame(wl.flame)
        hist = ctx.histogram(W, H)
        fh.warmup(19)
        d = fh.draw_to_bins(hist, 8)
        ovf1 = fh.sort_overflow()
        d += fh.draw_to_bins(hist, 8)
        out[mode] = (d, hist.download(), fh.particles(), ovf1, fh.sort_overflow())
    dq, hq, pq, _, _ = out[1]
    ds, hs, ps, o1, o2 = out[5]
    assert dq == ds > 0
    assert np.array_equal(pq.view(np.uint32), ps.view(np.uint32))
    assert np.array_equal(hq[..., 3], hs[..., 3])
    # A float32 running sum of n terms can drift by about n * 2^-24 of its
    # value; the quad path adds a hot bin's samples one by one, the sort path
    # adds per-tile partial sums, so for bins hit ~10^6 times (cfg3) the two
    # legitimately differ by ~1 %.
    # Tiles sum colours in fixed point (2^16 - 1 per unit, 2^15 - 1 for tiles
    # of more than 65532 records): up to 1 / (2 * 32767) per sample of
    # rounding.
    n = hq[..., 3:4].astype(np.float64)
    tol = 1e-3 + n / 65534.0 + np.abs(hq[..., :3]) * (2e-3 + n * 2.0 ** -23)
    assert (np.abs(hs[..., :3].astype(np.float64) - hq[..., :3]) <= tol).all()
    assert 0.0 <= o2 <= max(o1, 0.0) + 1e-9
    if cfg == "cfg2":
        assert o2 < 0.01


@pytest.mark.parametrize("size", [(1, 1), (3, 5), (67, 45)])
def test_tiny_and_ragged_histograms_every_mode(rf, gpu_available, size):
    """A zero-pass draw returns 0 and adds nothing; on 1x1, 3x5 and 67x45
    histograms every histogram mode (quad, wave aggregation, naive, LDS table,
    sort) adds the samples the quad path adds: equal alpha counts, `drawn`
    equal to their sum, colours within the modes' rounding.  AUTO's first draw
    of a flame runs its first pass as a launch of its own (the LDS probe), and
    every launch seeds its own streams, so its samples differ from a single
    launch's: there `drawn` equals the sum of its alpha counts."""
    from refrakt_amd import configs
    wl = configs.cfg1()
    W, H = size
    fl = wl.flame
    fl.size = [W, H]
    fl.scale = fl.scale * W / wl.width
    out = {}
    for mode in (1, 2, 3, 4, 5, 0):
        ctx = _ctx(rf, 1 << 14, 1, rf.RF_ENGINE_FAST, hist_mode=mode, seed=5, waves=8)
        fh = ctx.flame(fl)
        hist = ctx.histogram(W, H)
        fh.warmup(3)
        assert fh.draw_to_bins(hist, 0) == 0
        assert not hist.download().any()
        d = fh.draw_to_bins(hist, 3)
        out[mode] = (d, hist.download())
        fh.close()
        ctx.close()
    dq, hq = out[1]
    assert dq == int(hq[..., 3].astype(np.float64).sum()) > 0
    n = hq[..., 3:4].astype(np.float64)
    tol = 1e-3 + n / 65534.0 + np.abs(hq[..., :3]) * (2e-3 + n * 2.0 ** -23)
    for mode, (d, h) in out.items():
        assert d == int(h[..., 3].astype(np.float64).sum()), mode
        if mode == 0:
            continue
        assert d == dq, mode
        assert np.array_equal(h[..., 3], hq[..., 3]), mode
        assert (np.abs(h[..., :3].astype(np.float64) - hq[..., :3]) <= tol).all(), mode


@pytest.mark.parametrize("H", [8192, 8193])
def test_sort_path_at_its_size_limit(rf, gpu_available, H):
    """16384 x 8192 = 2^27 bins is the sort path's largest histogram (256
    buckets of 2^19 bins): sort-mode draws give the quad path's alpha counts
    exactly (first draw on uniform capacities, so hot buckets overflow into
    the atomic path; second on learned ones).  One row more and sort-mode
    draws fall back to quad atomics, with the same result."""
    import torch
    from refrakt_amd import configs
    wl = configs.cfg2()
    W = 16384
    fl = wl.flame
    fl.size = [W, H]
    fl.scale = fl.scale * W / wl.width
    res = {}
    for mode in (1, 5):
        ctx = _ctx(rf, 1 << 18, 1, rf.RF_ENGINE_FAST, hist_mode=mode, seed=3, waves=8)
        buf = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        hist = ctx.histogram(W, H, buf.data_ptr())
        fh = ctx.flame(fl)
        fh.warmup(5)
        d = fh.draw_to_bins(hist, 4)
        d += fh.draw_to_bins(hist, 4)
        res[mode] = (d, buf)
        if mode == 5:  # -1: no sort draw ran
            assert (fh.sort_overflow() >= 0.0) == (W * H <= 1 << 27)
        hist.close()
        fh.close()
        ctx.close()
    (dq, bq), (ds, bs) = res[1], res[5]
    assert dq == ds > 0
    assert torch.equal(bq[..., 3], bs[..., 3])
    assert int(bq[..., 3].double().sum()) == dq
    n = bq[..., 3:4].double()
    tol = 1e-3 + n / 65534.0 + bq[..., :3].double().abs() * (2e-3 + n * 2.0 ** -23)
    assert bool(((bs[..., :3].double() - bq[..., :3].double()).abs() <= tol).all())


def test_full_size_cfg2_conservation(rf, gpu_available):
    """cfg2 at its real size (7680x4320, 2^24 particles, 2^32 binned samples)."""
    from refrakt_amd import configs
    wl = configs.cfg2()
    ctx = _ctx(rf, wl.particles, wl.temporal_samples, rf.RF_ENGINE_FAST)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(wl.width, wl.height)
    fh.warmup(wl.warmup_passes)
    drawn = sum(fh.draw_to_bins(hist, wl.draw_passes) for _ in range(wl.draw_calls))
    g = hist.download()
    assert np.isfinite(g).all()
    assert g[..., 3].max() < 2 ** 24
    assert 0.9 * wl.binned_iterations <= drawn <= wl.binned_iterations
    assert int(g[..., 3].astype(np.float64).sum()) == drawn
    pos = fh.particles()
    assert np.isfinite(pos).all()


def test_rgba8_device_matches_host(rf, gpu_available):
    import torch
    ctx = _ctx(rf, 1 << 10, 1, rf.RF_ENGINE_FAST)
    rng = np.random.default_rng(9)
    img = rng.uniform(-0.2, 1.2, size=(67, 45, 4)).astype(np.float32)
    img[3, 4, 1] = np.nan
    a = torch.from_numpy(img).cuda()
    b = torch.empty((67, 45, 4), dtype=torch.uint8, device="cuda")
    ctx.rgba8(a.data_ptr(), b.data_ptr(), 45, 67)
    assert np.array_equal(b.cpu().numpy(), rf.to_rgba8(img))
