"""Flame programs: the fast engine compiled per flame structure (hiprtc).

CPU tests compile programs without a device (the source generator, every
variation's code under the runtime compiler, the build flags); GPU tests
check that a program computes the same bits as the precompiled interpreter:
particle positions after warmup and draws bit-exact, hit counts exact,
colour sums within 1e-5 relative (float atomics add in any order).
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from test_gpu_parity import _var_params

OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"


def _all_variation_flame(rf, seed=5, W=160, H=128):
    """A flame whose xforms together run every variation once (pre_xform ones
    next to a regular term), with post-affines on half of them and a final
    xform."""
    rng = np.random.default_rng(seed)
    names = rf.variation_names()
    pre = [n for n in names if rf.lib().rf_variation_flags(rf.variation_id(n)) & 2]
    reg = [n for n in names if n not in pre]
    groups = [reg[i:i + 6] for i in range(0, len(reg), 6)]
    prm_all = _var_params(None, rng)
    xs = []
    for k, g in enumerate(groups):
        vs = {n: float(rng.uniform(0.1, 0.4)) for n in g}
        if k < len(pre):
            vs[pre[k]] = 0.2
        prm = {}
        for n in vs:
            for p in rf.variation_params(n):
                prm[p] = prm_all.get(p, 0.5)
        th = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(0.4, 0.8)
        xs.append(rf.Xform(affine=[s * math.cos(th), s * math.sin(th), -s * math.sin(th), s * math.cos(th)]
                           + list(rng.uniform(-0.5, 0.5, 2)),
                           post=list(rng.uniform(-1, 1, 6)) if k % 2 else None, variations=vs, var_param=prm,
                           weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform()), color_speed=0.5,
                           opacity=1.0))
    fin = rf.Xform(affine=[0.9, 0.1, -0.1, 0.9, 0.05, -0.02], variations={"linear": 0.8, "spherical": 0.2},
                   color=0.3, color_speed=0.2, opacity=1.0)
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    pal[:, 1] = np.linspace(1, 0, 256)
    pal[:, 2] = 0.5
    covered = {n for x in xs for n in x.variations}
    assert covered == set(names)
    return rf.Flame(xforms=xs, final_xform=fin, palette=pal, size=[W, H], center=[0.0, 0.0], scale=W / 6.0)


def test_program_source_fixes_the_structure(rf):
    from refrakt_amd import configs
    fl = configs.cfg2().flame
    src = rf.program_source(fl, 8, 5)
    assert ('extern "C" __global__ __launch_bounds__(512) '
            '__attribute__((amdgpu_waves_per_eu(rfk::detail::fast_waves_per_eu<5>()))) void rf_prog_w8_h5') in src
    assert "k_fast_body<8, 5, 0, rfprog::Prog>" in src
    # one switch case per xform (the last as default), the final xform inline
    # (once with its term offset fixed, once for the batched renderer)
    assert src.count("return apply_fixed<") == len(fl.xforms) + 2
    assert "has_final(BP) { return true; }" in src
    assert "TERM_BEGIN_LOAD" not in src.split("struct BProg")[0]
    lin = rf.variation_id("linear")
    assert f"TermList<{lin}u>(), blk, FINAL" in src
    with pytest.raises(rf.RefraktError):
        rf.program_source(fl, 6, 5)
    with pytest.raises(rf.RefraktError):
        rf.program_source(fl, 8, 0)


def test_every_variation_compiles_in_a_program_without_packed_fp32(rf):
    """hiprtc compiles the kernel body with every variation's code (no device
    needed); the code object carries the program's entry point and no packed
    FP32 add/mul/fma (the build disables them, as for the precompiled kernels)."""
    fl = _all_variation_flame(rf)
    code, log, sec = rf.program_build(fl, 8, 5, "gfx950")
    assert len(code) > 4096, log
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "prog.co")
        with open(p, "wb") as f:
            f.write(code)
        dis = subprocess.run([OBJDUMP, "-d", "--no-show-raw-insn", p], capture_output=True, text=True, check=True)
    text = dis.stdout
    assert "<rf_prog_w8_h5>:" in text
    for op in ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32"):
        assert op not in text, op


def _run(rf, fl, programs, hist_mode, P=1 << 14, W=160, H=128, warm=3, draws=(4, 5), waves=8):
    ctx = rf.Context(P, 1, 16, seed=11, engine=rf.RF_ENGINE_FAST, hist_mode=hist_mode, block_waves=waves,
                     flame_programs=programs)
    fh = ctx.flame(fl)
    hist = ctx.histogram(W, H)
    fh.warmup(warm)
    pos0 = fh.particles().copy()
    drawn = [fh.draw_to_bins(hist, n) for n in draws]
    info = fh.program_info()
    out = (pos0, fh.particles().copy(), hist.download(), drawn, info, fh.hist_mode()[0])
    fh.close()
    ctx.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("hist_mode", [0, 1, 4, 5])
def test_program_matches_interpreter_bitwise(rf, gpu_available, hist_mode):
    fl = _all_variation_flame(rf)
    a = _run(rf, fl, True, hist_mode)
    b = _run(rf, fl, False, hist_mode)
    assert a[4][0] == 1 and b[4][0] == 0
    assert a[5] == b[5]
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert a[3] == b[3] and sum(a[3]) > 0
    assert np.array_equal(a[2][..., 3], b[2][..., 3])
    np.testing.assert_allclose(a[2][..., :3], b[2][..., :3], rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
def test_program_cfg2_flame_matches_interpreter(rf, gpu_available):
    from refrakt_amd import configs
    wl = configs.cfg2()
    W, H = 1920, 1080
    a = _run(rf, wl.flame, True, 0, P=1 << 18, W=W, H=H, warm=19, draws=(16,))
    b = _run(rf, wl.flame, False, 0, P=1 << 18, W=W, H=H, warm=19, draws=(16,))
    assert a[4][0] == 1 and b[4][0] == 0
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert a[3] == b[3]
    assert np.array_equal(a[2][..., 3], b[2][..., 3])
    np.testing.assert_allclose(a[2][..., :3], b[2][..., :3], rtol=1e-5, atol=1e-5)


def test_batch_program_compiles(rf):
    """The batched renderer's kernel body compiles as a program for the
    family structure of the cfg4 frames (mode 16 + 4: LDS table)."""
    from refrakt_amd import configs
    frames = configs.cfg4()
    big = max((f.flame for f in frames), key=lambda fl: len(fl.xforms))
    src = rf.program_source(big, 8, 16 + 4)
    assert "rf_batch_w8_h4" in src and "k_batch_body<8, 4, 11, rfprog::BProg>" in src
    # the batch kernel finds the final xform's terms through the Block
    assert ("TERM_BEGIN_LOAD>(" in src.split("struct BProg")[1]) == (big.final_xform is not None)
    code, log, sec = rf.program_build(big, 8, 16 + 4, "gfx950")
    assert len(code) > 4096, log
    # and with sort staging (mode 16 + 5)
    src = rf.program_source(big, 8, 16 + 5)
    assert "rf_batch_w8_h5" in src and "k_batch_body<8, 5, 0, rfprog::BProg>" in src
    code, log, sec = rf.program_build(big, 8, 16 + 5, "gfx950")
    assert len(code) > 4096, log
