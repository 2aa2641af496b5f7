"""Host-side logic that needs no GPU: parameter block, screen affine, validation."""
import math

import numpy as np
import pytest


def _flame(rf):
    x0 = rf.Xform(affine=[1, 2, 3, 4, 5, 6], variations={"spherical": 0.25, "linear": 0.75}, weight=1.0,
                  color=0.1, color_speed=0.2, opacity=0.9, rotation_frequency=1.0)
    x1 = rf.Xform(affine=[7, 8, 9, 10, 11, 12], post=[0.5, 0, 0, 0.5, 1, -1],
                  variations={"julian": 0.5, "blob": 0.5},
                  var_param={"julian_power": 3.0, "julian_dist": 1.5, "blob_high": 2.0, "blob_low": 0.5,
                             "blob_waves": 4.0},
                  weight=3.0, color=0.7, color_speed=0.4, opacity=1.0)
    fin = rf.Xform(affine=[1, 0, 0, 1, 0, 0], variations={"linear": 1.0}, weight=2.0, color=0.5, color_speed=0.0,
                   opacity=1.0)
    return rf.Flame(xforms=[x0, x1], final_xform=fin, size=[800, 600], center=[0.25, -0.5], scale=100.0,
                    rotate=30.0)


def test_param_block_layout(rf):
    """make_shader_buffer_map + copy_flame_data_to_buffer (src/flame.cpp:33-103):
    weight/sum, affine, [post], variation weights by name, params by name,
    color, color_speed, opacity, rotation_frequency; final xform last."""
    b = rf.param_block(_flame(rf))
    f32 = np.float32
    exp = [f32(1.0) / f32(4.0), 1, 2, 3, 4, 5, 6, 0.75, 0.25, 0.1, 0.2, f32(0.9), 1.0,
           f32(3.0) / f32(4.0), 7, 8, 9, 10, 11, 12, 0.5, 0, 0, 0.5, 1, -1,
           0.5, 0.5,                    # blob, julian (alphabetical)
           2.0, 0.5, 4.0, 1.5, 3.0,     # blob_high, blob_low, blob_waves, julian_dist, julian_power
           0.7, 0.4, 1.0, 0.0,
           f32(2.0) / f32(4.0), 1, 0, 0, 1, 0, 0, 1.0, 0.5, 0.0, 1.0, 0.0]
    assert b.tolist() == np.array(exp, np.float32).tolist()


def test_screen_affine_closed_form(rf):
    fl = _flame(rf)
    W, H = 1600, 1200
    ss = rf.screen_affine(fl, W, H).astype(np.float64)
    s = 100.0 * H / 600.0
    th = math.radians(30.0)
    # rotate_affine (src/flame.hpp:97-109) on s*I: a' = s cos, b' = s sin, c' = -s sin, d' = s cos
    a, b, c, d = s * math.cos(th), s * math.sin(th), -s * math.sin(th), s * math.cos(th)
    # T(W/2,H/2) S(s) R(30 deg) T(-center): x' = a x + c y + e
    e = W / 2 + a * -0.25 + c * 0.5
    f = H / 2 + b * -0.25 + d * 0.5
    np.testing.assert_allclose(ss, [a, b, c, d, e, f], rtol=1e-5, atol=1e-3)


def test_validate_rejects_bad_flames(rf):
    fl = _flame(rf)
    assert rf.validate(fl) == ""
    for x in fl.xforms:
        x.weight = 0.0
    assert "weights" in rf.validate(fl)
    fl2 = _flame(rf)
    fl2.xforms = []
    with pytest.raises(rf.RefraktError):
        fl2.to_c()


def test_configs_are_deterministic(rf):
    from refrakt_amd import configs
    a, b = configs.cfg2(), configs.cfg2()
    assert rf.param_block(a.flame).tolist() == rf.param_block(b.flame).tolist()
    assert a.binned_iterations == 1 << 32
    assert configs.cfg1().binned_iterations == 1 << 30
    assert configs.cfg3().binned_iterations == 1 << 34
    c2 = a.flame
    assert len(c2.xforms) == 8 and c2.final_xform is not None
    assert all(x.post is not None and 3 <= len(x.variations) <= 5 for x in c2.xforms)
    assert all(abs(sum(x.variations.values()) - 1) < 1e-6 for x in c2.xforms)


def _read_png(path):
    """Minimal PNG reader for 8-bit RGBA, non-interlaced files with filter-0 rows."""
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xffffffff == crc, typ
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype, comp, filt, inter = ihdr
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert (raw[:, 0] == 0).all()
    return raw[:, 1:].reshape(h, w, 4)


def test_png_round_trip_and_rgba8_conversion(rf, tmp_path):
    rng = np.random.default_rng(5)
    img = rng.uniform(-0.2, 1.2, size=(37, 53, 4)).astype(np.float32)
    img[0, 0, 0] = np.nan
    img[0, 1, :] = [0.5 / 255, 1.5 / 255, 254.5 / 255, 1.0]
    q = rf.to_rgba8(img)
    ref = np.floor(np.clip(np.nan_to_num(img, nan=0.0), 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(q, ref)
    assert q[0, 0, 0] == 0 and q[0, 1, 3] == 255
    path = str(tmp_path / "x.png")
    rf.write_png(path, q)
    assert np.array_equal(_read_png(path), q)


def _read_exr(path):
    """Minimal reader of the files rf_write_exr writes (single-part scanline,
    uncompressed): header attributes, offset table and (H, W, 4) RGBA pixels
    as raw uint16 (HALF) or float32 (FLOAT) values."""
    import struct
    b = open(path, "rb").read()
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01]) and b[4:8] == bytes([2, 0, 0, 0])
    p, attrs = 8, {}
    while b[p] != 0:
        e = b.index(0, p); name = b[p:e].decode(); p = e + 1
        e = b.index(0, p); typ = b[p:e].decode(); p = e + 1
        (n,) = struct.unpack_from("<i", b, p); p += 4
        attrs[name] = (typ, b[p:p + n]); p += n
    p += 1
    chans, q = [], attrs["channels"][1]
    while q[0] != 0:
        e = q.index(0); nm = q[:e].decode(); q = q[e + 1:]
        pt, lin, xs, ys = struct.unpack_from("<iB3xii", q); q = q[16:]
        chans.append((nm, pt, lin, xs, ys))
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack_from(f"<{H}Q", b, p)
    pt = chans[0][1]
    dt = np.dtype("<u2") if pt == 1 else np.dtype("<f4")
    img = np.zeros((H, W, 4), dt)
    for y in range(H):
        yy, n = struct.unpack_from("<ii", b, offs[y])
        assert yy == y and n == W * 4 * dt.itemsize
        line = np.frombuffer(b, dt, W * 4, offs[y] + 8).reshape(4, W)
        for k, (nm, _, _, _, _) in enumerate(chans):
            img[y, :, "RGBA".index(nm)] = line[k]
    return attrs, chans, img


def test_half_from_float_matches_ieee_round_to_nearest_even(rf):
    import ctypes
    L = rf.lib()
    edge = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 65519.99, 65520.0, 1e6, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25,
                     np.nextafter(np.float32(2.0 ** -25), np.float32(1)), 3 * 2.0 ** -25, 1.5 * 2.0 ** -24,
                     2.0 ** -14 - 2.0 ** -25, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 1e-10, np.inf, -np.inf],
                    np.float32)
    rng = np.random.default_rng(9)
    rnd = np.concatenate([edge, rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) **
                          rng.integers(-9, 6, 20000).astype(np.float32)]).astype(np.float32)
    got = np.array([L.rf_half_from_float(ctypes.c_float(float(v))) for v in rnd], np.uint16)
    with np.errstate(over="ignore"):  # values past 65504 become infinity, as they should
        ref = rnd.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, ref)
    nan = L.rf_half_from_float(ctypes.c_float(float("nan")))
    assert np.isnan(np.array([nan], np.uint16).view(np.float16)[0])


@pytest.mark.parametrize("half", [True, False])
def test_exr_round_trip(rf, tmp_path, half):
    rng = np.random.default_rng(3)
    img = (rng.uniform(-1, 4, size=(19, 31, 4)) * 10.0 ** rng.integers(-6, 5, size=(19, 31, 4))).astype(np.float32)
    img[0, 0] = [np.inf, -np.inf, 0.0, 70000.0]
    img[1, 2, 1] = np.nan
    path = str(tmp_path / "x.exr")
    rf.write_exr(path, img, half=half)
    attrs, chans, got = _read_exr(path)
    assert [c[0] for c in chans] == ["A", "B", "G", "R"]
    assert all(c[1] == (1 if half else 2) and c[3] == 1 and c[4] == 1 for c in chans)
    assert attrs["compression"] == ("compression", b"\x00")
    assert attrs["lineOrder"] == ("lineOrder", b"\x00")
    assert attrs["dataWindow"][1] == attrs["displayWindow"][1]
    if half:
        with np.errstate(over="ignore"):
            ref = img.astype(np.float16).view(np.uint16)
        nan = np.isnan(img)
        assert np.array_equal(got[~nan], ref[~nan])
        assert np.isnan(got.view(np.float16)[nan]).all()
    else:
        assert np.array_equal(got.view(np.uint32), img.view(np.uint32))
