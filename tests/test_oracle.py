"""Pins the CPU oracle against published vectors, closed forms and invariants.

The oracle (oracle/oracle.c) restates the reference's algorithm; these checks
are chosen so that a plausible slip (a transposed affine coefficient, atan2
arguments swapped, a dropped normalisation, a wrong rotation constant) fails.
"""
import math

import numpy as np
import pytest

from refrakt_amd.api import Flame, Xform


# ------------------------------------------------------------------ RNGs
def test_philox4x32_10_known_answers(orc):
    """Random123 kat_vectors for philox4x32-10."""
    L = orc.lib()
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    out = np.zeros(4, np.uint32)
    for ctr, key, exp in cases:
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        L.orc_philox4x32_10(orc.up(c), orc.up(k), orc.up(out))
        assert tuple(int(v) for v in out) == exp


def _jsf_py(seed):
    """Bob Jenkins' small PRNG (burtleburtle.net/bob/rand/smallprng.html), 32-bit
    version, in Python integers: e = a - rot(b,27); a = b ^ rot(c,17); b = c + d;
    c = d + e; d = e + a; raninit: a = 0xf1ea5eed, b = c = d = seed, 20 rounds."""
    M = 0xffffffff
    rot = lambda x, k: ((x << k) | (x >> (32 - k))) & M
    s = [0xf1ea5eed, seed, seed, seed]

    def step():
        a, b, c, d = s
        e = (a - rot(b, 27)) & M
        a = b ^ rot(c, 17)
        b = (c + d) & M
        c = (d + e) & M
        d = (e + a) & M
        s[:] = [a, b, c, d]
        return e
    for _ in range(20):
        step()
    return s, step


def test_jsf32_matches_published_definition(orc):
    import ctypes as C
    L = orc.lib()
    for seed in [0, 1, 7, 123456789, 0xffffffff]:
        st = orc.Jsf()
        L.orc_jsf_seed(C.byref(st), seed)
        s, step = _jsf_py(seed)
        assert (st.a, st.b, st.c, st.d) == tuple(s)
        for _ in range(50):
            e = step()
            a = L.orc_jsf_next_a(C.byref(st))
            assert a == s[0] and (st.a, st.b, st.c, st.d) == tuple(s)
            # the CPU variant returns d = e + a of the same step (util.hpp:87)
            assert (e + a) & 0xffffffff == st.d


def test_jsf32_randf_is_u_times_2_pow_minus_32(orc):
    import ctypes as C
    L = orc.lib()
    st = orc.Jsf()
    L.orc_jsf_seed(C.byref(st), 42)
    st2 = orc.Jsf(st.a, st.b, st.c, st.d)
    for _ in range(1000):
        f = L.orc_jsf_randf(C.byref(st))
        u = L.orc_jsf_next_a(C.byref(st2))
        assert f == np.float32(np.float32(u) * np.float32(2.0 ** -32))
        assert 0.0 <= f <= 1.0


def test_hammersley_brute_force(orc):
    for count in [4096, 768, 7, 256]:
        M = count if count % 2 == 0 else 1 << (count - 1).bit_length()
        r = M.bit_length() - 1
        pts = orc.hammersley(count)
        inv = np.float32(1.0) / np.float32(M)   # the reference scales by float(1/max)
        for i in range(count):
            low = i & ((1 << r) - 1)
            rev = int(format(low, f"0{r}b")[::-1], 2) if r else 0
            assert pts[i, 0] == np.float32(float(np.float32(i) * inv) * 2.0 - 1.0)
            assert pts[i, 1] == np.float32(float(np.float32(rev) * inv) * 2.0 - 1.0)
            if M & (M - 1) == 0:                 # power of two: exact radical inverse
                assert pts[i, 1] == np.float32(2.0 * rev / M - 1.0)


# ----------------------------------------------------------- xform selection
def _flame(weights):
    return Flame(xforms=[Xform(weight=w, variations={"linear": 1.0}, opacity=1.0) for w in weights])


def test_select_boundaries_and_frequencies(orc):
    import ctypes as C
    L = orc.lib()
    fl = orc.flame_to_c(_flame([1.0, 3.0, 4.0]))
    cdf = np.zeros(32, np.float32)
    L.orc_cdf(C.byref(fl), orc.fp(cdf))
    assert cdf[0] == np.float32(1 / 8) and cdf[1] == np.float32(0.5) and cdf[2] == np.float32(1.0)
    assert L.orc_select(C.byref(fl), 0.0) == 0
    assert L.orc_select(C.byref(fl), float(np.float32(0.125))) == 0        # cdf >= u is inclusive
    assert L.orc_select(C.byref(fl), float(np.nextafter(np.float32(0.125), np.float32(1)))) == 1
    assert L.orc_select(C.byref(fl), 1.0) == 2                                # last xform unconditional
    one = orc.flame_to_c(_flame([2.5]))
    assert L.orc_select(C.byref(one), 1.0) == 0
    rng = np.random.default_rng(1)
    u = rng.random(200000).astype(np.float32)
    counts = np.bincount([L.orc_select(C.byref(fl), float(v)) for v in u], minlength=3)
    exp = np.array([1, 3, 4]) / 8 * len(u)
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < 13.8  # 99.9% quantile, 2 dof


# ------------------------------------------------------------- variations
HAND = [  # (name, x, y, rands, params, expected (x, y)); a = atan2(x, y), phi = atan2(y, x)
    ("linear", 0.3, -0.7, None, None, (0.3, -0.7)),
    ("sinusoidal", math.pi / 2, 0.0, None, None, (1.0, 0.0)),
    ("spherical", 3.0, 4.0, None, None, (3 / 25, 4 / 25)),
    ("swirl", 1.0, 0.0, None, None, (math.sin(1), math.cos(1))),
    ("horseshoe", 3.0, 4.0, None, None, (-7 / 5, 24 / 5)),
    ("polar", 1.0, 0.0, None, None, (0.5, 0.0)),           # a = atan2(1, 0) = pi/2
    ("polar", 0.0, 2.0, None, None, (0.0, 1.0)),
    ("handkerchief", 0.0, 2.0, None, None, (2 * math.sin(2), 2 * math.cos(2))),
    ("heart", 0.0, 2.0, None, None, (0.0, -2.0)),
    ("disc", 1.0, 0.0, None, None, (0.0, -0.5)),
    ("spiral", 3.0, 4.0, None, None, ((0.8 + math.sin(5)) / 5, (0.6 - math.cos(5)) / 5)),
    ("julia", 0.0, 4.0, [0.2], None, (2.0, 0.0)),
    ("julia", 0.0, 4.0, [0.8], None, (-2.0, 0.0)),
    ("bubble", 2.0, 0.0, None, None, (1.0, 0.0)),
    ("fisheye", 3.0, 4.0, None, None, (2 / 6 * 4, 2 / 6 * 3)),
    ("eyefish", 3.0, 4.0, None, None, (2 / 6 * 3, 2 / 6 * 4)),
    ("exponential", 1.0, 0.5, None, None, (0.0, 1.0)),     # e^0 (cos pi/2, sin pi/2)
    ("cosine", 0.0, 0.0, None, None, (1.0, 0.0)),
    ("hyperbolic", 3.0, 4.0, None, None, (0.6 / 5, 0.8 * 5)),
    ("diamond", 3.0, 4.0, None, None, (0.6 * math.cos(5), 0.8 * math.sin(5))),
    ("bent", -1.0, -1.0, None, None, (-2.0, -0.5)),
    ("cylinder", math.pi / 6, 0.25, None, None, (0.5, 0.25)),
    ("exp", 0.0, math.pi, None, None, (-1.0, 0.0)),
    ("log", 0.0, math.e, None, None, (1.0, math.pi / 2)),
    ("rectangles", 0.25, -0.25, None, {"rectangles_x": 1.0, "rectangles_y": 1.0}, (0.75, -0.75)),
    ("blur", 0.0, 0.0, [0.5, 0.25], None, (0.0, 0.5)),
    ("square", 9.0, 9.0, [0.75, 0.25], None, (0.25, -0.25)),
    ("mobius", 0.5, 0.0, None, {"mobius_re_a": 1, "mobius_re_d": 1}, (0.5, 0.0)),  # identity map
    ("curl", 1.0, 0.0, None, {"curl_c1": 1.0}, (0.5, 0.0)),                         # z / (1 + z)
]


@pytest.mark.parametrize("name,x,y,rands,params,exp", HAND)
def test_variation_hand_computed(orc, name, x, y, rands, params, exp):
    ox, oy = orc.var_eval(name, x, y, 1.0, params, rands=(rands or []) + [0] * (16 - len(rands or [])))
    assert ox == pytest.approx(exp[0], abs=1e-9) and oy == pytest.approx(exp[1], abs=1e-9)


def test_variation_invariants(orc):
    rng = np.random.default_rng(4)
    for x, y in rng.uniform(-2, 2, (200, 2)):
        r = math.hypot(x, y)
        sx, sy = orc.var_eval("swirl", x, y)
        assert math.hypot(sx, sy) == pytest.approx(r, rel=1e-12)            # swirl rotates
        hx, hy = orc.var_eval("horseshoe", x, y)
        assert math.hypot(hx, hy) == pytest.approx(r * r / (r + 1e-10), rel=1e-9)
        jx, jy = orc.var_eval("julia", x, y, rands=[rng.random()] + [0] * 15)
        assert jx * jx + jy * jy == pytest.approx(r, rel=1e-12)            # |julia|^2 = r
        px, py = orc.var_eval("spherical", *orc.var_eval("spherical", x, y))
        # inversion is an involution up to the EPS = 1e-10 in the denominator
        # (relative error eps/r^2 from the first inversion, eps*r^2 from the second)
        assert (px, py) == pytest.approx((x, y), rel=2e-10 * (1 / (r * r) + r * r) + 1e-12)
        dx, dy = orc.var_eval("disc", x, y)
        assert math.hypot(dx, dy) == pytest.approx(abs(math.atan2(x, y)) / math.pi, rel=1e-12)
        kx, ky = orc.var_eval("heart", x, y)
        assert math.hypot(kx, ky) == pytest.approx(r, rel=1e-12)


# ------------------------------------------------- chaos game (compat, f32)
def _affine_flame(affines, weights=None):
    xs = [Xform(affine=list(a), variations={"linear": 1.0}, weight=(weights[i] if weights else 1.0),
                color=0.5, color_speed=0.0, opacity=1.0) for i, a in enumerate(affines)]
    pal = np.ones((256, 4), np.float32)
    return Flame(xforms=xs, palette=pal, size=[64, 64], scale=8.0)


def _run_compat(orc, fl, T, ppt, passes, seed_rows=3):
    P = T * ppt
    fc = orc.flame_to_c(fl)
    rng = np.random.default_rng(seed_rows)
    R = 4
    shuf = np.stack([rng.permutation(ppt).astype(np.uint32) for _ in range(R)])
    rs = orc.jsf_seed_states(P)
    samples = orc.hammersley(ppt)
    a = np.zeros((P, 4), np.float32)
    b = np.zeros((P, 4), np.float32)
    orc.compat_pass(fc, T, ppt, samples, a, shuf, rs, 0, 1, 1, 1, 1, 0)
    for i in range(passes):
        src, dst = (a, b) if i % 2 == 0 else (b, a)
        orc.compat_pass(fc, T, ppt, src, dst, shuf, rs, i % R, (i + 1) % R, 1, 1, 0, 0)
    return (b if passes % 2 else a), shuf, rs


def test_compat_converges_to_closed_form_fixed_point(orc):
    """x' = a x + c y + e, y' = b x + d y + f (flam3 coefs a b c d e f): the
    particles of a single contraction end at the fixed point of M p + t."""
    a, b, c, d, e, f = 0.5, -0.1, 0.2, 0.4, 1.0, -0.5
    pos, _, _ = _run_compat(orc, _affine_flame([(a, b, c, d, e, f)]), 2, 512, 60)
    M = np.array([[a, c], [b, d]])
    p = np.linalg.solve(np.eye(2) - M, np.array([e, f]))
    assert np.abs(pos[:, :2] - p).max() < 1e-5


def test_compat_identity_pass_is_a_permutation(orc):
    fl = _affine_flame([(1, 0, 0, 1, 0, 0)])
    T, ppt = 2, 512
    P = T * ppt
    fc = orc.flame_to_c(fl)
    rng = np.random.default_rng(0)
    src = rng.uniform(-1, 1, (P, 4)).astype(np.float32)
    src[:, 3] = 0
    dst = np.zeros_like(src)
    shuf = np.stack([rng.permutation(ppt).astype(np.uint32) for _ in range(2)])
    rs = orc.jsf_seed_states(P)
    orc.compat_pass(fc, T, ppt, src, dst, shuf, rs, 0, 1, 1, 1, 0, 0)
    for y in range(T):
        s = src[y * ppt:(y + 1) * ppt, :2]
        o = dst[y * ppt:(y + 1) * ppt, :2]
        assert np.array_equal(np.sort(s.view(np.uint64).ravel()), np.sort(o.view(np.uint64).ravel()))
        # o[out_row[g]] = s[in_row[g]]
        assert np.array_equal(o[shuf[1]], s[shuf[0]])


def test_compat_workgroup_shares_one_xform(orc):
    """Two constant maps (x -> 0 and x -> 1) with weights 1:3: after one pass
    each 256-lane workgroup lands on a single value, and the share of
    workgroups on 1 is close to 3/4 (flame.glsl:51-53)."""
    fl = _affine_flame([(0, 0, 0, 0, 0, 0), (0, 0, 0, 0, 1, 0)], weights=[1.0, 3.0])
    T, ppt = 8, 4096
    pos, shuf, _ = _run_compat(orc, fl, T, ppt, 1)
    # the last pass wrote lane gx of sample y to slot y*ppt + shuf[1][gx]
    by_lane = np.concatenate([pos[y * ppt + shuf[1].astype(np.int64), 0] for y in range(T)])
    groups = by_lane.reshape(-1, 256)
    assert np.all(groups.min(axis=1) == groups.max(axis=1))
    frac = (groups[:, 0] == 1.0).mean()
    assert abs(frac - 0.75) < 0.08


def test_compat_draw_counts_match_positions(orc):
    """Binning: every in-view sample adds exactly one hit at (H-1-floor(sy))*W + floor(sx)."""
    fl = _affine_flame([(0.5, 0, 0, 0.5, 0.5, 0.5), (0.5, 0, 0, 0.5, -0.5, 0.5), (0.5, 0, 0, 0.5, 0, -0.5)])
    fl.scale = 20.0
    T, ppt, W, H = 1, 1024, 64, 48
    pos, shuf, rs = _run_compat(orc, fl, T, ppt, 12)
    fc = orc.flame_to_c(fl)
    ss = orc.screen_affine(fl, W, H)
    out = np.zeros_like(pos)
    hits = np.zeros((H, W), np.uint32)
    drawn = orc.compat_pass(fc, T, ppt, pos, out, shuf, rs, 0, 0, 1, 0, 0, 1, W, H, ss, None, hits)
    sx = np.floor(ss[0] * out[:, 0].astype(np.float64) + ss[2] * out[:, 1] + ss[4])
    sy = np.floor(ss[1] * out[:, 0].astype(np.float64) + ss[3] * out[:, 1] + ss[5])
    ok = (sx >= 0) & (sx < W) & (sy >= 0) & (sy < H)
    ref = np.zeros((H, W), np.uint32)
    np.add.at(ref, ((H - 1 - sy[ok]).astype(int), sx[ok].astype(int)), 1)
    assert drawn == ok.sum() == hits.sum()
    assert np.array_equal(hits, ref)


@pytest.mark.parametrize("threads", [1, 3])
def test_ref_passes_mt_equals_compat_pass(orc, threads):
    """orc_ref_passes_mt (the CPU baseline: the reference's passes spread over
    threads, float histogram) computes orc_compat_pass's positions, jsf32
    states and drawn counts bit for bit; on one thread its histogram holds
    exactly the compat pass's hit counts; on several, adds to one bin may race
    as in the shader (flame.glsl:82-84), so alpha may only fall short."""
    from refrakt_amd.api import Flame, Xform
    xs = [Xform(affine=[0.5, 0.1, -0.2, 0.6, 0.3 * k - 0.3, 0.2], variations={"linear": 0.7, "spherical": 0.3},
                weight=1.0 + k, color=k / 3.0, color_speed=0.5, opacity=1.0) for k in range(3)]
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    fl = Flame(xforms=xs, palette=pal, size=[96, 80], scale=30.0)
    T, ppt, W, H = 4, 2048, 96, 80
    P = T * ppt
    fc = orc.flame_to_c(fl)
    rng = np.random.default_rng(4)
    R = 6
    shuf = np.stack([rng.permutation(ppt).astype(np.uint32) for _ in range(R)])
    rows = rng.integers(0, R, size=2 * 10).astype(np.uint32)
    ss = orc.screen_affine(fl, W, H)
    samples = orc.hammersley(ppt)
    # single-threaded oracle pass by pass
    rs1 = orc.jsf_seed_states(P)
    a, b = np.zeros((P, 4), np.float32), np.zeros((P, 4), np.float32)
    orc.compat_pass(fc, T, ppt, samples, a, shuf, rs1, rows[0], rows[1], 1, 1, 1, 0)
    bufs = [a, b]
    for i in range(5):
        orc.compat_pass(fc, T, ppt, bufs[i % 2], bufs[(i + 1) % 2], shuf, rs1, rows[2 + 2 * i], rows[3 + 2 * i],
                        1, 1, 0, 0)
    hits = np.zeros((H, W), np.uint32)
    d1 = 0
    for i in range(4):
        d1 += orc.compat_pass(fc, T, ppt, bufs[(5 + i) % 2], bufs[(6 + i) % 2], shuf, rs1, rows[12 + 2 * i],
                              rows[13 + 2 * i], 1, 0, 0, 1, W, H, ss, None, hits)
    final1 = bufs[(5 + 4) % 2]
    # the multithreaded restatement
    rs2 = orc.jsf_seed_states(P)
    A, B = np.zeros((P, 4), np.float32), np.zeros((P, 4), np.float32)
    orc.ref_passes_mt(fc, T, ppt, samples, A, shuf, rs2, rows[:2], 1, 1, 1, 1, 0, threads=threads)
    orc.ref_passes_mt(fc, T, ppt, A, B, shuf, rs2, rows[2:12], 5, 1, 1, 0, 0, threads=threads)
    bins = np.zeros((H, W, 4), np.float32)
    d2 = orc.ref_passes_mt(fc, T, ppt, B, A, shuf, rs2, rows[12:20], 4, 1, 0, 0, 1, W, H, ss, bins, threads)
    final2 = B   # 5 warmup passes end in B; 4 draw passes B->A->B->A->B
    assert d1 == d2 > 0
    assert np.array_equal(rs1, rs2)
    assert np.array_equal(final1.view(np.uint32), final2.view(np.uint32))
    if threads == 1:
        assert np.array_equal(bins[..., 3].astype(np.uint32), hits)
    else:
        assert (bins[..., 3] <= hits + 0.5).all() and bins[..., 3].sum() >= 0.99 * d2


# --------------------------------------------------------- chaos game (f64)
def test_chaos_f64_sierpinski_mass(orc):
    """Three maps p -> (p + v_i)/2: the attractor is the Sierpinski triangle,
    each corner sub-triangle holds 1/3 of the mass, the central hole none."""
    v = [(-1.0, -1.0), (1.0, -1.0), (0.0, 1.0)]
    fl = _affine_flame([(0.5, 0, 0, 0.5, vx / 2, vy / 2) for vx, vy in v])
    W = H = 256
    fl.size, fl.scale = [W, H], W / 2.2
    bins, n = orc.chaos_f64(fl, W, H, 20000, 30, 50, seed=5, threads=4)
    assert n == 20000 * 50
    a = bins[..., 3]
    ss = orc.screen_affine(fl, W, H).astype(np.float64)
    iy, ix = np.nonzero(a >= 0)
    cx = ix + 0.5
    cy = H - 1 - iy + 0.5
    det = ss[0] * ss[3] - ss[1] * ss[2]
    fx = (ss[3] * (cx - ss[4]) - ss[2] * (cy - ss[5])) / det
    fy = (-ss[1] * (cx - ss[4]) + ss[0] * (cy - ss[5])) / det
    m = a[iy, ix]
    assert m.sum() == n
    # central hole: triangle with corners (0,-1), (-0.5, 0), (0.5, 0), shrunk by 2 pixels
    eps = 2.5 * 2.2 / W
    inside = (fy > -1 + eps) & (fy < -eps) & (np.abs(fx) < (fy + 1) / 2 - eps)
    assert m[inside].sum() == 0
    left, right, top = (fy < 0) & (fx < 0) & ~inside, (fy < 0) & (fx > 0) & ~inside, fy >= 0
    for part in (left, right, top):
        assert abs(m[part].sum() / n - 1 / 3) < 0.02


# ------------------------------------------------------------------- DE
def test_density_identity_and_clip(orc):
    W, H = 20, 15
    h = np.zeros((H, W, 4), np.float32)
    h[2, 5] = [1, 2, 3, 4]
    h[7, 0] = [9, 9, 9, 9]      # column 0: sprite centre outside the clip volume
    out = orc.density(h, 0, 0, 0.4)
    exp = np.zeros((H, W, 4))
    exp[H - 1 - 2, 4] = [1, 2, 3, 4]
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("r", [3, 9, 30])
def test_density_stamp_mass_closed_form(orc, r):
    """Sum of the Epanechnikov stamp weights -> S^2/(4 r^2) (the continuous
    integral of (1 - |2p/S|^2) * 2/(pi r^2) over the disc |p| <= S/2, S = 2r+1)."""
    W = H = 2 * r + 9
    h = np.zeros((H, W, 4), np.float32)
    h[H // 2, W // 2] = [1, 1, 1, 1]             # a = 1 -> radius = R
    out = orc.density(h, r, 0, 0.4, fp32=False)
    S = 2 * r + 1
    assert out[..., 3].sum() == pytest.approx(S * S / (4 * r * r), rel=3.0 / r)
    # the stamp is centred on texel (px - 1, H - 1 - py) and spans radius r
    nz = np.argwhere(out[..., 3] > 0)
    cy, cx = H - 1 - H // 2, W // 2 - 1
    assert nz[:, 0].min() == cy - r + (1 if r > 3 else 0) or nz[:, 0].min() >= cy - r
    assert np.abs(nz[:, 0] - cy).max() <= r and np.abs(nz[:, 1] - cx).max() <= r


def _gl_point_splat(h, R, rmin, curve):
    """The density pass restated from the OpenGL 4.6 rules, independently of
    the oracle's closed form (float64): one point per bin at object position
    pos - offset, offset = (hw, -hw), hw = 1/(4r + 2) (density_vert.glsl:27-60)
    through ortho(0, W, H, 0) and viewport (0, 0, W, H) (main.cpp:497-511);
    a point whose centre leaves the clip volume is culled; a size-S point
    covers the fragments whose centres lie in the S x S square around its
    window position (spec 14.4); gl_PointCoord = (0.5 + (xc - xw) / S,
    0.5 - (yc - yw) / S) (origin UPPER_LEFT); density_frag.glsl:9-19 weights.
    Output rows are window rows (row 0 at the bottom)."""
    H, W = h.shape[:2]
    out = np.zeros((H, W, 4))
    for py in range(H):
        for px in range(W):
            c = h[py, px].astype(np.float64)
            if c[3] == 0.0:
                continue
            r = max(rmin, min(R, int(R / c[3] ** curve)))
            hw = 1.0 / (4.0 * r + 2.0)
            ox, oy = px - hw, py + hw                    # object position
            xw, yw = ox, H - oy                           # window position (ortho with y flipped)
            if not (0.0 <= xw <= W and 0.0 <= yw <= H):
                continue
            S = 2 * r + 1
            for fy in range(int(np.floor(yw - S / 2)) - 1, int(np.ceil(yw + S / 2)) + 1):
                for fx in range(int(np.floor(xw - S / 2)) - 1, int(np.ceil(xw + S / 2)) + 1):
                    xc, yc = fx + 0.5, fy + 0.5
                    if not (abs(xc - xw) <= S / 2 and abs(yc - yw) <= S / 2):
                        continue
                    if not (0 <= fx < W and 0 <= fy < H):
                        continue
                    if r == 0:
                        out[fy, fx] += c
                        continue
                    s, t = 0.5 + (xc - xw) / S, 0.5 - (yc - yw) / S
                    nx, ny = 2.0 * (s + hw) - 1.0, 2.0 * (t - hw) - 1.0
                    d = nx * nx + ny * ny
                    if d > 1.0:
                        continue
                    out[fy, fx] += c * (1.0 - d) * (0.63661977236 / (r * r))
    return out


@pytest.mark.parametrize("R,rmin", [(0, 0), (1, 0), (3, 0), (9, 2)])
def test_density_stamp_matches_gl_point_rasterization(orc, R, rmin):
    """Pins the oracle's stamp geometry (sub-pixel offset, centre texel,
    chord ends, weights, clipping at the borders and the culled column 0)
    against the point rasterisation restated from the GL rules."""
    rng = np.random.default_rng(40 + R)
    W, H = 2 * R + 13, 2 * R + 11
    h = np.zeros((H, W, 4), np.float32)
    for py, px in [(0, 0), (0, 1), (H - 1, W - 1), (H // 2, W // 2), (1, W - 2), (H - 2, 3)]:
        a = float(rng.choice([1.0, 3.0, 50.0]))
        h[py, px] = [rng.uniform(0, a), rng.uniform(0, a), rng.uniform(0, a), a]
    got = orc.density(h, R, rmin, 0.4, fp32=False)
    ref = _gl_point_splat(h, R, rmin, 0.4)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert (got[..., 3] > 0).sum() == (ref[..., 3] > 0).sum() > 0


def test_density_radius_shrinks_with_density(orc):
    W = H = 41
    widths = []
    for a in [1.0, 10.0, 1000.0, 1e6]:
        h = np.zeros((H, W, 4), np.float32)
        h[20, 20] = [a, a, a, a]
        out = orc.density(h, 9, 0, 0.4)
        widths.append(int((out[..., 3] > 0).sum()))
    assert widths == sorted(widths, reverse=True) and widths[-1] == 1


# --------------------------------------------------------------- tonemap
def test_tonemap_closed_form(orc):
    k = 1e-4
    a = (10.0 ** 2 - 1.0) / k                   # log10(1 + a k) = 2
    b = 4.0
    ls = 0.5 * b * 2 / a                        # 4 / a
    img = np.array([[[a / 8, a / 16, a / 32, a]]], np.float64)   # rgb * ls = 0.5, 0.25, 0.125 ; a * ls = 4
    # vibrancy 1, gamma 2: rgb * ls * sqrt(4) / 4 = rgb * ls / 2
    out = orc.tonemap(img, 2.0, b, 1.0, k)
    assert out[0, 0, :3] == pytest.approx([0.25, 0.125, 0.0625]) and out[0, 0, 3] == 1.0
    # vibrancy 0, gamma 2: sqrt(rgb * ls)
    out = orc.tonemap(img, 2.0, b, 0.0, k)
    assert out[0, 0, :3] == pytest.approx([math.sqrt(0.5), 0.5, math.sqrt(0.125)])
    # empty texel -> black, alpha 1; saturation clamps to 1
    z = orc.tonemap(np.zeros((1, 1, 4)), 2.2, 4.0, 1.0, k)
    assert list(z[0, 0]) == [0, 0, 0, 1]
    s = orc.tonemap(np.array([[[1e9, 0, 0, 1e3]]]), 1.0, 100.0, 1.0, k)
    assert s[0, 0, 0] == 1.0
