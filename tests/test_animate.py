"""Temporal-sample animation (shaders/templates/animate.tpl.glsl:18-96).

CPU: the oracle's orc_animate against the closed form.  Sample i of T rotates
every xform's linear part by theta = 0.31415926535 * dt * rotation_frequency,
dt = (i - T/2) / (T/2) * tss, multiplying [a c; b d] by the rotation
[cos -sin; sin cos] on the right (a' = a cos + c sin, c' = c cos - a sin, as
flame::rotate_affine, src/flame.hpp:97-109).  Pinned by a float64 matrix
product (a left multiplication or a sign slip fails it), the middle sample
being the identity bit for bit, the determinant being preserved, opposite
samples rotating by opposite angles, T = 1 meaning no rotation, and every
other field being copied.

GPU: the compat engine with rotating xforms against orc_animate + the
oracle's compat pass, bit-exact (positions, jsf32 states, hit counts) for an
affine/linear flame, colour sums within the float summation bound n 2^-24 |sum|
of a bin of n samples; the fast engine at T = 1 with rotating xforms equals the
same flame without rotation bit for bit.
"""
import math

import numpy as np
import pytest

RPS = 0.31415926535


def _general_flame(rf, seed, n=3, final=True, W=192, H=160, rot=(1.0, -2.0, 0.0), rot_final=0.5):
    """Linear flame whose affines are general matrices (shear + scale), so
    left and right multiplication by a rotation give different results."""
    rng = np.random.default_rng(seed)
    xs = []
    for k in range(n):
        th = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(0.35, 0.7)
        sh = rng.uniform(-0.3, 0.3)
        a, b, c, d = s * math.cos(th), s * math.sin(th), -s * math.sin(th) + sh, s * math.cos(th)
        xs.append(rf.Xform(affine=[a, b, c, d] + list(rng.uniform(-0.8, 0.8, 2)),
                           post=list(rng.uniform(-1, 1, 6)) if k == 1 else None,
                           variations={"linear": float(rng.uniform(0.7, 1.1))}, weight=float(rng.uniform(0.2, 1)),
                           color=float(rng.uniform()), color_speed=0.5, opacity=1.0,
                           rotation_frequency=float(rot[k % len(rot)])))
    fin = None
    if final:
        fin = rf.Xform(affine=[0.9, 0.15, -0.1, 0.8, 0.05, -0.02], variations={"linear": 1.0}, color=0.3,
                       color_speed=0.2, opacity=1.0, rotation_frequency=rot_final)
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    pal[:, 1] = np.linspace(1, 0, 256)
    pal[:, 2] = 0.5
    return rf.Flame(xforms=xs, final_xform=fin, palette=pal, size=[W, H], center=[0.0, 0.0], scale=W / 5.0)


def _aff(ox):
    return np.array(ox.aff[:], np.float64)


def _rot_closed(aff, theta):
    M = np.array([[aff[0], aff[2]], [aff[1], aff[3]]], np.float64)
    R = np.array([[math.cos(theta), -math.sin(theta)], [math.sin(theta), math.cos(theta)]])
    Mr = M @ R
    return np.array([Mr[0, 0], Mr[1, 0], Mr[0, 1], Mr[1, 1], aff[4], aff[5]])


def _theta(i, T, tss, rot):
    half = T // 2
    dt = (i - half) / half * tss if half else 0.0
    return RPS * dt * rot


# ------------------------------------------------------------------ CPU pins
@pytest.mark.parametrize("T", [32, 64, 2, 33])
def test_oracle_animate_closed_form(rf, orc, T):
    fl = _general_flame(rf, 3)
    fc = orc.flame_to_c(fl)
    tss = 1.5
    for i in range(T):
        o = orc.animate(fc, i, T, tss)
        assert o.n_xforms == fc.n_xforms and o.has_final == fc.has_final
        pairs = [(fc.xf[k], o.xf[k]) for k in range(fc.n_xforms)] + [(fc.fin, o.fin)]
        for src, dst in pairs:
            th = _theta(i, T, tss, src.rotation_frequency)
            want = _rot_closed(_aff(src), th)
            np.testing.assert_allclose(_aff(dst), want, rtol=0, atol=2e-6)
            # determinant of the linear part is preserved (det R = 1)
            a = _aff(src)
            b = _aff(dst)
            assert abs((b[0] * b[3] - b[1] * b[2]) - (a[0] * a[3] - a[1] * a[2])) < 3e-6
            # only the linear part changes
            assert _aff(dst)[4:].tolist() == _aff(src)[4:].tolist()
            assert list(dst.post) == list(src.post) and dst.has_post == src.has_post
            assert (dst.weight, dst.color, dst.color_speed, dst.opacity, dst.n_terms) == \
                   (src.weight, src.color, src.color_speed, src.opacity, src.n_terms)
            assert all(dst.terms[t].weight == src.terms[t].weight for t in range(src.n_terms))


def test_oracle_animate_middle_and_symmetry(rf, orc):
    fl = _general_flame(rf, 5)
    fc = orc.flame_to_c(fl)
    T, tss = 32, 0.9
    mid = orc.animate(fc, T // 2, T, tss)
    for k in range(fc.n_xforms):
        assert np.array_equal(_aff(mid.xf[k]).astype(np.float32).view(np.uint32),
                              _aff(fc.xf[k]).astype(np.float32).view(np.uint32)), "middle sample must be unrotated"
    # samples half - j and half + j rotate by opposite angles: undoing one with
    # R(+2 theta) gives the other
    for j in (1, 5, 15):
        lo = orc.animate(fc, T // 2 - j, T, tss)
        hi = orc.animate(fc, T // 2 + j, T, tss)
        for k in range(fc.n_xforms):
            th = _theta(T // 2 + j, T, tss, fc.xf[k].rotation_frequency)
            np.testing.assert_allclose(_rot_closed(_aff(lo.xf[k]), 2 * th), _aff(hi.xf[k]), atol=3e-6)
    # a xform with rotation_frequency 0 is never rotated
    zero = [k for k in range(fc.n_xforms) if fc.xf[k].rotation_frequency == 0.0]
    assert zero
    for i in (0, 7, 31):
        o = orc.animate(fc, i, T, tss)
        for k in zero:
            assert _aff(o.xf[k]).tolist() == _aff(fc.xf[k]).tolist()


def test_oracle_animate_single_sample_is_unrotated(rf, orc):
    fl = _general_flame(rf, 7)
    fc = orc.flame_to_c(fl)
    o = orc.animate(fc, 0, 1, 2.0)
    for k in range(fc.n_xforms):
        assert _aff(o.xf[k]).tolist() == _aff(fc.xf[k]).tolist()
    assert _aff(o.fin).tolist() == _aff(fc.fin).tolist()


def test_rotate_affine_matches_reference_formula():
    """configs.rotate_affine restates flame::rotate_affine (src/flame.hpp:97-109):
    a rotation by deg degrees multiplied on the right; 90 degrees maps
    [a c; b d] to [c -a; d -b]."""
    from refrakt_amd import configs
    a = [0.3, -0.2, 0.7, 0.4, 0.1, 0.2]
    r = configs.rotate_affine(a, 90.0)
    np.testing.assert_allclose(r, [0.7, 0.4, -0.3, 0.2, 0.1, 0.2], atol=1e-6)
    th = 0.01745329251 * 33.0
    np.testing.assert_allclose(configs.rotate_affine(a, 33.0), _rot_closed(a, th), atol=1e-6)


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("T", [32, 1])
def test_compat_animated_bit_exact(rf, orc, gpu_available, T):
    P, R, W, H = 1 << 15, 16, 192, 160
    ppt = P // T
    tss = 1.5
    fl = _general_flame(rf, 11, W=W, H=H)
    ctx = rf.Context(P, T, R, seed=7, engine=rf.RF_ENGINE_COMPAT, hist_mode=1, block_waves=4)
    rows_rng = np.random.default_rng(9)
    warm, nd = 3, 4
    wrows = rows_rng.integers(0, R, size=2 * (warm + 1)).astype(np.uint32)
    drows = rows_rng.integers(0, R, size=2 * nd).astype(np.uint32)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    print(f"compat animated: T={T} P={P} tss={tss}", flush=True)
    fh.warmup(warm, tss_width=tss, rows=wrows)
    drawn = fh.draw_to_bins(hist, nd, rows=drows)
    g_hist = hist.download()
    g_pos = fh.particles()
    g_rs = ctx.rng_states()
    shuf = ctx.shuffle_rows()

    fc = orc.flame_to_c(fl)
    per = [orc.animate(fc, i, T, tss) for i in range(T)]
    rs = orc.jsf_seed_states(P)
    bufs = [np.zeros((P, 4), np.float32), np.zeros((P, 4), np.float32)]
    kw = dict(flames_per_sample=per)
    orc.compat_pass(fc, T, ppt, orc.hammersley(ppt), bufs[0], shuf, rs, wrows[0], wrows[1], 1, 1, 1, 0, **kw)
    for i in range(warm):
        orc.compat_pass(fc, T, ppt, bufs[i % 2], bufs[(i + 1) % 2], shuf, rs, wrows[2 + 2 * i], wrows[3 + 2 * i],
                        1, 1, 0, 0, **kw)
    if warm % 2:
        bufs = [bufs[1], bufs[0]]
    ss = orc.screen_affine(fl, W, H)
    bins = np.zeros((H, W, 4), np.float64)
    hits = np.zeros((H, W), np.uint32)
    o_drawn = 0
    for i in range(nd):
        o_drawn += orc.compat_pass(fc, T, ppt, bufs[i % 2], bufs[(i + 1) % 2], shuf, rs, drows[2 * i],
                                   drows[2 * i + 1], 1, 0, 0, 1, W, H, ss, bins, hits, **kw)
    final = bufs[nd % 2]
    assert drawn == o_drawn > 0
    assert np.array_equal(g_pos.view(np.uint32), final.view(np.uint32)), "particle positions differ"
    assert np.array_equal(g_rs, rs), "jsf32 states differ"
    assert np.array_equal(g_hist[..., 3].astype(np.uint32), hits), "hit counts (bin indices) differ"
    # float atomics add a bin's n samples in any order: |error| <= n 2^-24 |sum|
    tol = 1e-6 + hits[..., None] * 2.0 ** -24 * np.abs(bins[..., :3])
    assert (np.abs(g_hist[..., :3] - bins[..., :3]) <= tol).all()
    if T > 1:
        # the animation is visible: the unrotated flame bins differently
        ref = np.zeros((H, W), np.uint32)
        b2 = [np.zeros((P, 4), np.float32), np.zeros((P, 4), np.float32)]
        rs2 = orc.jsf_seed_states(P)
        orc.compat_pass(fc, T, ppt, orc.hammersley(ppt), b2[0], shuf, rs2, wrows[0], wrows[1], 1, 1, 1, 0)
        for i in range(warm):
            orc.compat_pass(fc, T, ppt, b2[i % 2], b2[(i + 1) % 2], shuf, rs2, wrows[2 + 2 * i],
                            wrows[3 + 2 * i], 1, 1, 0, 0)
        if warm % 2:
            b2 = [b2[1], b2[0]]
        for i in range(nd):
            orc.compat_pass(fc, T, ppt, b2[i % 2], b2[(i + 1) % 2], shuf, rs2, drows[2 * i], drows[2 * i + 1],
                            1, 0, 0, 1, W, H, ss, None, ref)
        assert not np.array_equal(ref, hits)


@pytest.mark.gpu
def test_fast_single_sample_rotation_is_identity(rf, gpu_available):
    """T = 1 with rotating xforms: the one temporal sample is the centre, so
    the fast engine bins exactly what the same flame without rotation bins."""
    P, W, H = 1 << 14, 192, 160
    fl_rot = _general_flame(rf, 13, W=W, H=H)
    fl_still = _general_flame(rf, 13, W=W, H=H, rot=(0.0,), rot_final=0.0)
    out = []
    for fl in (fl_rot, fl_still):
        ctx = rf.Context(P, 1, 16, seed=21, engine=rf.RF_ENGINE_FAST, hist_mode=1, block_waves=4)
        fh = ctx.flame(fl)
        hist = ctx.histogram(W, H)
        fh.warmup(5)
        d = fh.draw_to_bins(hist, 3)
        out.append((d, hist.download()[..., 3].copy()))
        hist.close()
        ctx.close()
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])
