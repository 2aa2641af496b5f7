"""Pins of the oracle's variations beyond the twelve of the benchmark set
(variations.yaml:64-550): values worked out by hand at chosen points where
the formula collapses to a closed form, and, for the complex-analytic
variations, Python's cmath.  Each case names what a plausible transcription
slip would change: a swapped sin/cos pair (`.yx`), a parameter or affine
coefficient read from the wrong slot, a dropped term, a wrong branch.

Notation: a = atan2(x, y), phi = atan2(y, x) (variations.yaml:4-5); the
affine coefficients c10, c11, c20, c21 are flam3's c, d, e, f (index
column * 2 + row, src/variation_table.cpp:52-57).  Random draws come from
a table (`rands`, read in order); parameters and affines travel as float32,
hence the 1e-6 tolerance where they enter.
"""
import cmath
import math

import pytest

PI = math.pi
S2 = math.sqrt(2.0)

# (name, x, y, rands, params, affine, weight, expected, tolerance)
CASES = [
    # ex: a = atan2(0, 1) = 0, r = 1: (sin^3 1 + cos^3 1, sin^3 1 - cos^3 1)
    ("ex", 0.0, 1.0, None, None, None, 1.0,
     (math.sin(1) ** 3 + math.cos(1) ** 3, math.sin(1) ** 3 - math.cos(1) ** 3), 1e-12),
    # waves reads c10 = c and c20 = e for x, c11 = d and c21 = f for y
    ("waves", 2 * PI, PI / 2, None, None, (1.0, 0.0, 0.5, 0.25, 1.0, 2.0), 1.0, (2 * PI + 0.5, PI / 2 + 0.25), 1e-8),
    # popcorn: tan(3 pi / 12) = 1, tan 0 = 0; e scales x, f scales y
    ("popcorn", 0.0, PI / 12, None, None, (1.0, 0.0, 0.0, 1.0, 2.0, 3.0), 1.0, (2 * math.sin(1.0), PI / 12), 1e-7),
    # power: r^sina (cosa, sina), sina = x / r, cosa = y / r
    ("power", 3.0, 4.0, None, None, None, 1.0, (5 ** 0.6 * 0.8, 5 ** 0.6 * 0.6), 1e-12),
    ("power", 0.0, 2.0, None, None, None, 1.0, (1.0, 0.0), 1e-12),
    # rings (dx = e^2 = 1): r = 0.5 -> m = 0.5, r = 3.5 -> m = -0.5; times (cosa, sina)
    ("rings", 0.0, 0.5, None, None, (1.0, 0.0, 0.0, 1.0, 1.0, 0.0), 1.0, (0.5, 0.0), 1e-9),
    ("rings", 0.0, 3.5, None, None, (1.0, 0.0, 0.0, 1.0, 1.0, 0.0), 1.0, (-0.5, 0.0), 1e-8),
    # fan (dx = pi, offset f = 0): a = pi/4 -> 3pi/4; a = 3pi/4 -> pi/4; result r (cos, sin)
    ("fan", 1.0, 1.0, None, None, (1.0, 0.0, 0.0, 1.0, 1.0, 0.0), 1.0, (-1.0, 1.0), 1e-8),
    ("fan", 1.0, -1.0, None, None, (1.0, 0.0, 0.0, 1.0, 1.0, 0.0), 1.0, (1.0, 1.0), 1e-8),
    # blob: a = 0 -> m = r (low + high) / 2 along (sina, cosa) = (0, 1); a = pi/2, waves 1 -> m = r high
    ("blob", 0.0, 2.0, None, {"blob_high": 1.25, "blob_low": 0.5, "blob_waves": 3.0}, None, 1.0, (0.0, 1.75), 1e-9),
    ("blob", 1.0, 0.0, None, {"blob_high": 1.25, "blob_low": 0.5, "blob_waves": 1.0}, None, 1.0, (1.25, 0.0), 1e-9),
    # pdj: (sin(a y) - cos(b x), sin(c x) - cos(d y))
    ("pdj", PI / 6, PI / 2, None, {"pdj_a": 1.0, "pdj_b": 2.0, "pdj_c": 3.0, "pdj_d": 4.0}, None, 1.0, (0.5, 0.0), 1e-9),
    ("pdj", 0.0, 0.0, None, {"pdj_a": 1.0, "pdj_b": 2.0, "pdj_c": 3.0, "pdj_d": 4.0}, None, 1.0, (-1.0, -1.0), 1e-12),
    # fan2 (dx = pi, fan2_y = 0): a = 0 -> pi/2, a = 3pi/4 -> pi/4; result r (sin, cos), no swap
    ("fan2", 0.0, 1.0, None, {"fan2_x": 1.0, "fan2_y": 0.0}, None, 1.0, (1.0, 0.0), 1e-8),
    ("fan2", 1.0, -1.0, None, {"fan2_x": 1.0, "fan2_y": 0.0}, None, 1.0, (1.0, 1.0), 1e-8),
    # rings2 (dx = 1): r = 0.5 stays, r = 3.5 -> -0.5; along (sina, cosa)
    ("rings2", 0.0, 0.5, None, {"rings2_val": 1.0}, None, 1.0, (0.0, 0.5), 1e-9),
    ("rings2", 0.0, 3.5, None, {"rings2_val": 1.0}, None, 1.0, (0.0, -0.5), 1e-8),
    # perspective: angle 0 is the identity; angle pi/2, dist 2 at (1, 1): (2, 0) / (2 - 1)
    ("perspective", 0.3, -0.7, None, {"perspective_angle": 0.0, "perspective_dist": 2.0}, None, 1.0, (0.3, -0.7), 1e-7),
    ("perspective", 1.0, 1.0, None, {"perspective_angle": PI / 2, "perspective_dist": 2.0}, None, 1.0, (2.0, 0.0), 1e-6),
    # noise: v * m * (cos t, sin t), t = 2 pi u
    ("noise", 2.0, 3.0, [0.5, 0.0], None, None, 1.0, (1.0, 0.0), 1e-9),
    ("noise", 2.0, 3.0, [1.0, 0.25], None, None, 1.0, (0.0, 3.0), 1e-7),
    # juliascope power 2, k = 1: t = (2 pi - phi) / 2 -> (-sqrt 2, sqrt 2) at z = 4i
    ("juliascope", 0.0, 4.0, [0.75], {"juliascope_power": 2.0, "juliascope_dist": 1.0}, None, 1.0, (-S2, S2), 1e-9),
    ("juliascope", 0.0, 4.0, [0.2], {"juliascope_power": 2.0, "juliascope_dist": 1.0}, None, 1.0, (S2, S2), 1e-9),
    # gaussian_blur and pre_blur: (sum of 4 draws - 2) (cos, sin)(2 pi u5)
    ("gaussian_blur", 9.0, 9.0, [1.0, 1.0, 1.0, 0.5, 0.0], None, None, 1.0, (1.5, 0.0), 1e-9),
    ("gaussian_blur", 9.0, 9.0, [0.0, 0.0, 0.0, 0.0, 0.25], None, None, 1.0, (0.0, -2.0), 1e-7),
    ("pre_blur", 9.0, 9.0, [1.0, 1.0, 1.0, 0.5, 0.0], None, None, 1.0, (1.5, 0.0), 1e-9),
    # radial_blur angle 0: zoom only, g = w (sum - 2) -> (g x, g y); angle 1: pure spin by g
    ("radial_blur", 0.6, -0.2, [1.0, 1.0, 1.0, 1.0], {"radial_blur_angle": 0.0}, None, 1.0, (1.2, -0.4), 1e-9),
    ("radial_blur", 1.0, 0.0, [0.75, 0.75, 0.75, 0.75], {"radial_blur_angle": 1.0}, None, 1.0,
     (math.cos(1.0) - 1.0, math.sin(1.0)), 1e-7),
    # pie: slice trunc(0.3 * 4 + 0.5) = 1 -> t = pi/2; (m sin t, m cos t)
    ("pie", 5.0, 5.0, [1.0, 0.3, 0.0], {"pie_slices": 4.0, "pie_rotation": 0.0, "pie_thickness": 0.0}, None, 1.0,
     (1.0, 0.0), 1e-7),
    # ngon: corners 0, circle 1, power 0 is the identity; corners 1 at (2, 1): m = 1/cos(atan 1/2) - 1
    ("ngon", 0.3, -0.7, None, {"ngon_power": 0.0, "ngon_sides": 5.0, "ngon_corners": 0.0, "ngon_circle": 1.0}, None,
     1.0, (0.3, -0.7), 1e-9),
    ("ngon", 2.0, 1.0, None, {"ngon_power": 0.0, "ngon_sides": 4.0, "ngon_corners": 1.0, "ngon_circle": 0.0}, None,
     1.0, (2 * (math.sqrt(5) / 2 - 1), math.sqrt(5) / 2 - 1), 1e-8),
    # arch: t = pi/4 -> (sin t, sin^2 t / cos t)
    ("arch", 9.0, 9.0, [0.25], None, None, 1.0, (S2 / 2, S2 / 2), 1e-9),
    # tangent: (sin x / cos y, tan y)
    ("tangent", PI / 2, PI / 4, None, None, None, 1.0, (S2, 1.0), 1e-12),
    # rays: tan(pi/4) * w / rsq * (cos x, sin y) at (0, pi/2)
    ("rays", 0.0, PI / 2, [0.25], None, None, 1.0, (4 / PI ** 2, 4 / PI ** 2), 1e-9),
    # blade: t = u w r; (x (cos + sin), x (cos - sin))
    ("blade", 3.0, 4.0, [0.0], None, None, 1.0, (3.0, 3.0), 1e-12),
    ("blade", 3.0, 4.0, [PI / 10], None, None, 1.0, (3.0, -3.0), 1e-6),
    # secant2: cos(w r) = 0.5 -> 1/cr - 1; cos = -1 -> 1/cr + 1
    ("secant2", 0.0, PI / 3, None, None, None, 1.0, (0.0, 1.0), 1e-12),
    ("secant2", 0.0, PI, None, None, None, 1.0, (0.0, 0.0), 1e-12),
    # twintrian (flam3 form): t = pi/2 -> diff = log10 1 + 0 = 0 -> (0, -x pi)
    ("twintrian", 2.0, 0.0, [PI / 4], None, None, 1.0, (0.0, -2 * PI), 1e-6),
    # cross: v / |x^2 - y^2|
    ("cross", 2.0, 1.0, None, None, None, 1.0, (2 / 3, 1 / 3), 1e-9),
    # disc2 twist 0: (sin t, cos t) a / pi with t = rot pi (x + y); at (1, 0) a / pi = 1/2
    ("disc2", 1.0, 0.0, None, {"disc2_rot": 0.5, "disc2_twist": 0.0}, None, 1.0, (0.5, 0.0), 1e-9),
    # disc2 twist pi: add (cos pi - 1) k to x and sin(pi) k to y, k = 1 + pi + 2 pi
    ("disc2", 1.0, 0.0, None, {"disc2_rot": 0.5, "disc2_twist": PI}, None, 1.0,
     (0.5 * (1.0 - 2.0 * (1.0 + 3.0 * PI)), 0.0), 1e-5),
    # super_shape: m 0, n = 2 2 2 -> identity; n = 1 1 1 -> v / sqrt 2 (|cos| + |sin| at pi/4)
    ("super_shape", 0.3, -0.7, [0.0], {"super_shape_n1": 2.0, "super_shape_n2": 2.0, "super_shape_n3": 2.0,
                                       "super_shape_m": 0.0, "super_shape_holes": 0.0, "super_shape_rnd": 0.0}, None,
     1.0, (0.3, -0.7), 1e-7),
    ("super_shape", 0.3, -0.7, [0.0], {"super_shape_n1": 1.0, "super_shape_n2": 1.0, "super_shape_n3": 1.0,
                                       "super_shape_m": 0.0, "super_shape_holes": 0.0, "super_shape_rnd": 0.0}, None,
     1.0, (0.3 / S2, -0.7 / S2), 1e-7),
    # super_shape rnd 1 with u = 0.5, holes 0.25: v (0.5 - 0.25) / r
    ("super_shape", 3.0, 4.0, [0.5], {"super_shape_n1": 2.0, "super_shape_n2": 2.0, "super_shape_n3": 2.0,
                                      "super_shape_m": 0.0, "super_shape_holes": 0.25, "super_shape_rnd": 1.0}, None,
     1.0, (0.15, 0.2), 1e-7),
    # flower: phi = 0 -> v (u - holes) / r
    ("flower", 2.0, 0.0, [1.0], {"flower_holes": 0.0, "flower_petals": 5.0}, None, 1.0, (1.0, 0.0), 1e-9),
    ("flower", 0.0, 2.0, [1.0], {"flower_holes": 0.0, "flower_petals": 2.0}, None, 1.0, (0.0, -1.0), 1e-9),
    # conic: x = 0 -> v (u - holes) e / r
    ("conic", 0.0, 2.0, [1.0], {"conic_holes": 0.0, "conic_eccentricity": 0.5}, None, 1.0, (0.0, 0.5), 1e-9),
    # parabola: r = pi/2 -> (h u1, 0)
    ("parabola", 0.0, PI / 2, [0.5, 0.7], {"parabola_height": 0.8, "parabola_width": 1.2}, None, 1.0, (0.4, 0.0),
     1e-7),
    # bent2 (flam3 form): negative coordinates scaled by bent2_x / bent2_y
    ("bent2", -1.0, -2.0, None, {"bent2_x": 3.0, "bent2_y": 0.5}, None, 1.0, (-3.0, -1.0), 1e-9),
    ("bent2", 1.0, 2.0, None, {"bent2_x": 3.0, "bent2_y": 0.5}, None, 1.0, (1.0, 2.0), 1e-12),
    # bipolar (flam3 form): the origin maps to (0, 1); (1/2, 0) to (ln 9 / (2 pi), 1)
    ("bipolar", 0.0, 0.0, None, {"bipolar_shift": 0.0}, None, 1.0, (0.0, 1.0), 1e-12),
    ("bipolar", 0.5, 0.0, None, {"bipolar_shift": 0.0}, None, 1.0, (math.log(9.0) / (2 * PI), 1.0), 1e-12),
    # boarders: u >= 0.75 halves the offset from the nearest lattice point
    ("boarders", 1.2, -0.6, [0.8], None, None, 1.0, (1.1, -0.8), 1e-9),
    ("boarders", 1.2, -0.6, [0.1], None, None, 1.0, (1.225, -0.55), 1e-9),
    # butterfly at (2, 1): sqrt(2 / 8) = 1/2
    ("butterfly", 2.0, 1.0, None, None, None, 1.0, (1.3029400317411197, 1.3029400317411197), 1e-9),
    # cell (flam3 form), size 1: quadrant-dependent cell renumbering
    ("cell", 0.25, 0.75, None, {"cell_size": 1.0}, None, 1.0, (0.25, 0.75), 1e-12),
    ("cell", 1.5, 2.25, None, {"cell_size": 1.0}, None, 1.0, (2.5, 4.25), 1e-12),
    ("cell", -0.5, 0.5, None, {"cell_size": 1.0}, None, 1.0, (1.5, 0.5), 1e-12),
    ("cell", -0.5, -0.5, None, {"cell_size": 1.0}, None, 1.0, (1.5, 1.5), 1e-12),
    ("cell", 0.5, -0.5, None, {"cell_size": 1.0}, None, 1.0, (0.5, 1.5), 1e-12),
    # curve at the origin adds the amplitudes; at (1, 2) their Gaussians
    ("curve", 0.0, 0.0, None, {"curve_xlength": 2.0, "curve_ylength": 1.0, "curve_xamp": 0.5, "curve_yamp": -0.25},
     None, 1.0, (0.5, -0.25), 1e-9),
    ("curve", 1.0, 2.0, None, {"curve_xlength": 2.0, "curve_ylength": 1.0, "curve_xamp": 0.5, "curve_yamp": -0.25},
     None, 1.0, (1 + 0.5 / math.e, 2 - 0.25 / math.e), 1e-9),
    # foci: (sinh x, sin y) / (cosh x - cos y); at (ln 2, pi/2): (0.75, 1) / 1.25
    ("foci", math.log(2.0), PI / 2, None, None, None, 1.0, (0.6, 0.8), 1e-12),
    # oscope (flam3 form): below the wave y flips sign
    ("oscope", 0.0, 0.5, None, {"oscope_separation": 0.0, "oscope_frequency": 1.0, "oscope_amplitude": 1.0,
                                "oscope_damping": 0.0}, None, 1.0, (0.0, -0.5), 1e-12),
    ("oscope", 0.5, 0.5, None, {"oscope_separation": 0.0, "oscope_frequency": 1.0, "oscope_amplitude": 1.0,
                                "oscope_damping": 0.0}, None, 1.0, (0.5, 0.5), 1e-12),
    # loonie (w = 1): inside the unit circle v sqrt(1/rsq - 1), outside v
    ("loonie", 0.6, 0.0, None, None, None, 1.0, (0.8, 0.0), 1e-9),
    ("loonie", 2.0, 0.0, None, None, None, 1.0, (2.0, 0.0), 1e-12),
    # waves2: (x + sx sin(y fx), y + sy sin(x fy))
    ("waves2", PI / 2, PI / 4, None, {"waves2_scalex": 0.5, "waves2_scaley": 0.25, "waves2_freqx": 2.0,
                                      "waves2_freqy": 1.0}, None, 1.0, (PI / 2 + 0.5, PI / 4 + 0.25), 1e-9),
    # auger (freq 1, scale 0.5, sym 0.4, weight 0.6) at (pi/2, 0): y += 0.6 * 0.25, x keeps
    ("auger", PI / 2, 0.0, None, {"auger_freq": 1.0, "auger_scale": 0.5, "auger_sym": 0.4, "auger_weight": 0.6},
     None, 1.0, (PI / 2, 0.15), 1e-7),
    # flux (w = 1, spread 0): at i the two angles give t = pi/4 and m = 2; at 2, m = 2 sqrt 3
    ("flux", 0.0, 1.0, None, {"flux_spread": 0.0}, None, 1.0, (S2, S2), 1e-12),
    ("flux", 2.0, 0.0, None, {"flux_spread": 0.0}, None, 1.0, (2 * math.sqrt(3.0), 0.0), 1e-12),
]


@pytest.mark.parametrize("name,x,y,rands,params,affine,w,exp,tol", CASES,
                         ids=[f"{c[0]}-{k}" for k, c in enumerate(CASES)])
def test_variation_closed_form(orc, name, x, y, rands, params, affine, w, exp, tol):
    kw = {}
    if affine is not None:
        kw["affine"] = affine
    ox, oy = orc.var_eval(name, x, y, w, params, rands=(rands or []) + [0.0] * (16 - len(rands or [])), **kw)
    assert ox == pytest.approx(exp[0], abs=tol) and oy == pytest.approx(exp[1], abs=tol), (ox, oy)


def test_edisc_elliptic_closed_form(orc):
    """edisc (flam3 form) at the origin and at i: xmax = (sqrt(t + 2x) +
    sqrt(t - 2x)) / 2 with t = r^2 + 1, a1 = ln(xmax + sqrt(xmax - 1)),
    a2 = -acos(x / xmax); y > 0 negates sin a1."""
    ox, oy = orc.var_eval("edisc", 0.0, 0.0)
    assert ox == pytest.approx(math.cosh(PI / 2) / 11.57034632, abs=1e-12) and oy == pytest.approx(0.0, abs=1e-12)
    a1 = math.log(S2 + math.sqrt(S2 - 1.0))
    ox, oy = orc.var_eval("edisc", 0.0, 1.0)
    assert ox == pytest.approx(math.cosh(PI / 2) * math.cos(a1) / 11.57034632, abs=1e-12)
    assert oy == pytest.approx(math.sinh(PI / 2) * math.sin(a1) / 11.57034632, abs=1e-12)


def test_complex_analytic_variations_against_cmath(orc):
    """sin z, sinh z (variations.yaml:470, 476) are the complex functions;
    the reference's cos (:473) and coth (:500) are the conjugates of cmath's
    (cos x cosh y, +sin x sinh y; (sinh 2x, +sin 2y) / (cosh 2x - cos 2y));
    julian with power p, dist 1 and draw k is the k-th p-th root; cpow
    (:413) with draw 0 is exp((r + i im) / p * log z)."""
    import numpy as np
    rng = np.random.default_rng(11)
    for x, y in rng.uniform(-1.5, 1.5, (40, 2)):
        z = complex(x, y)
        for name, want in (("sin", cmath.sin(z)), ("sinh", cmath.sinh(z)),
                           ("cos", cmath.cos(z).conjugate()), ("coth", (1.0 / cmath.tanh(z)).conjugate())):
            ox, oy = orc.var_eval(name, x, y)
            assert complex(ox, oy) == pytest.approx(want, rel=1e-9, abs=1e-12), name
        for p, u in ((2.0, 0.2), (2.0, 0.7), (3.0, 0.5), (5.0, 0.95)):
            k = int(abs(p) * np.float32(u))
            want = abs(z) ** (1.0 / p) * cmath.exp(1j * (cmath.phase(z) + 2 * PI * k) / p)
            ox, oy = orc.var_eval("julian", x, y, params={"julian_power": p, "julian_dist": 1.0},
                                  rands=[u] + [0.0] * 15)
            assert complex(ox, oy) == pytest.approx(want, rel=1e-6, abs=1e-9), ("julian", p, u)
            assert want ** p == pytest.approx(z, rel=1e-6, abs=1e-9)   # a p-th root of z
        for pw, pr, pi_ in ((1.0, 2.0, 0.0), (2.0, 1.0, 0.0), (1.0, 1.0, 0.5), (3.0, 2.0, -0.25)):
            want = cmath.exp(complex(pr, pi_) / pw * cmath.log(z))
            ox, oy = orc.var_eval("cpow", x, y, params={"cpow_power": pw, "cpow_r": pr, "cpow_i": pi_},
                                  rands=[0.0] * 16)
            assert complex(ox, oy) == pytest.approx(want, rel=1e-6, abs=1e-9), ("cpow", pw, pr, pi_)
