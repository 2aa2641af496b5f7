"""flam3 reader (flame::load_flame, src/flame.cpp:160-226) on a fixture of this repository."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_reader_fields(rf):
    f = rf.load_flame(os.path.join(GOLD, "reader_fixture.flam3"))
    assert f.size == [640, 480]
    assert f.center == pytest.approx([0.25, -0.125])
    assert f.scale == pytest.approx(150.5) and f.rotate == pytest.approx(30)
    assert f.brightness == pytest.approx(12.5) and f.gamma == pytest.approx(2.5) and f.vibrancy == pytest.approx(0.75)
    assert f.estimator_radius == 7 and f.estimator_curve == pytest.approx(0.5)
    assert f.estimator_min == 0          # the reference reads "estimator_min", flam3 writes "estimator_minimum"
    assert len(f.xforms) == 2 and f.final_xform is not None
    x0, x1 = f.xforms
    assert x0.affine == pytest.approx([0.5, -0.25, 0.125, 0.75, 0.1, -0.2]) and x0.post is None
    assert x0.variations == pytest.approx({"linear": 0.6, "julian": 0.4})
    assert x0.var_param["julian_power"] == 3 and x0.var_param["julian_dist"] == -1.5
    assert x0.rotation_frequency == 1.0 and x1.rotation_frequency == 0.0     # animate > 0 -> 1
    assert x1.post == pytest.approx([0.9, 0.1, -0.1, 0.9, 0.05, 0.0])
    assert x1.var_param["rectangles_x"] == pytest.approx(0.2) and x1.var_param["rectangles_y"] == 0.0
    assert x1.opacity == 0.5 and x1.weight == 1.5 and x1.color_speed == 0.5
    # palette: std::stoi(token) / 256, alpha 1; unspecified entries stay zero
    assert f.palette[0].tolist() == [255 / 256, 128 / 256, 0, 1]
    assert f.palette[255].tolist() == [1 / 256, 2 / 256, 3 / 256, 1]
    assert f.palette[1].tolist() == [0, 0, 0, 0]


def test_reader_rejects_unknown_attribute(rf):
    txt = open(os.path.join(GOLD, "reader_fixture.flam3")).read().replace('linear="0.6"', 'no_such_var="0.6"')
    with pytest.raises(rf.RefraktError, match="Unknown attribute no_such_var"):
        rf.parse_flam3(txt)


def test_reader_rejects_malformed_xml(rf):
    with pytest.raises(rf.RefraktError):
        rf.parse_flam3('<flame size="1 1"><xform weight="1" linear="1" coefs="1 0 0 1 0 0"></flame>')
    with pytest.raises(rf.RefraktError):
        rf.parse_flam3("<flame/>")     # no xforms


def test_reader_roundtrip_through_the_abi(rf):
    f = rf.load_flame(os.path.join(GOLD, "reader_fixture.flam3"))
    d = f.to_c()
    g = rf.Flame.from_c(d)
    assert rf.param_block(f).tolist() == rf.param_block(g).tolist()
    assert np.array_equal(f.palette, g.palette)
