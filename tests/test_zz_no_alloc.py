"""rf_draw_to_bins_noalloc against rf_draw_to_bins, through the C ABI.

The allocation-free draw keeps its counters in a static device array shared
by every call on the GPU.  Two calls in a row, the second on a smaller
histogram with fewer passes (so counters left over from the first would show
in its drawn count), must give what the allocating entry point gives in the
same mode on a twin context: drawn counts equal, hit counts bit-exact,
colour sums within 2 n 2^-24 |sum| for a bin of n samples (float atomics add
in any order).  AUTO, whose choice for this flame is the sort path, runs the
LDS table and equals a fixed LDS-mode draw.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flame(rf, seed=41):
    rng = np.random.default_rng(seed)
    xs = []
    for k in range(4):
        th = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(0.4, 0.8)
        xs.append(rf.Xform(affine=[s * math.cos(th), s * math.sin(th), -s * math.sin(th), s * math.cos(th)]
                           + list(rng.uniform(-0.7, 0.7, 2)),
                           variations={"linear": 0.7, "spherical": 0.3} if k % 2 else {"linear": 1.0},
                           weight=float(rng.uniform(0.3, 1.0)), color=float(rng.uniform()), color_speed=0.5,
                           opacity=1.0))
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    pal[:, 2] = np.linspace(1, 0, 256)
    return rf.Flame(xforms=xs, palette=pal, size=[320, 256], center=[0.0, 0.0], scale=320 / 4.5)


def _same(a, b):
    assert np.array_equal(a[..., 3], b[..., 3]), "hit counts differ"
    tol = 1e-6 + 2.0 * a[..., 3:4] * 2.0 ** -24 * np.abs(a[..., :3])
    assert (np.abs(a[..., :3] - b[..., :3]) <= tol).all(), "colour sums differ"


@pytest.mark.parametrize("mode_ref,mode_noalloc", [(1, 1), (4, 4), (4, 0)])
def test_noalloc_draw_equals_draw_to_bins(rf, gpu_available, mode_ref, mode_noalloc):
    P = 1 << 16
    fl = _flame(rf)
    print(f"noalloc: reference mode {mode_ref}, no-alloc mode {mode_noalloc}", flush=True)
    ctx_a = rf.Context(P, 1, 16, seed=9, engine=rf.RF_ENGINE_FAST, hist_mode=mode_ref, block_waves=8)
    ctx_b = rf.Context(P, 1, 16, seed=9, engine=rf.RF_ENGINE_FAST, hist_mode=mode_noalloc, block_waves=8)
    fa, fb = ctx_a.flame(fl), ctx_b.flame(fl)
    fa.warmup(4)
    fb.warmup(4)
    for (W, H, n) in ((320, 256, 5), (160, 96, 2)):
        ha, hb = ctx_a.histogram(W, H), ctx_b.histogram(W, H)
        da = fa.draw_to_bins(ha, n)
        db = fb.draw_to_bins_noalloc(hb, n)
        assert da == db > 0, (W, H, n, da, db)
        a, b = ha.download(), hb.download()
        assert int(b[..., 3].sum(dtype=np.float64)) == db
        _same(a, b)
        ha.close()
        hb.close()
    assert np.array_equal(fa.particles().view(np.uint32), fb.particles().view(np.uint32))
    for o in (fa, fb):
        o.close()
    ctx_a.close()
    ctx_b.close()


def test_noalloc_rejects_compat_engine_and_cold_flame(rf, gpu_available):
    fl = _flame(rf)
    ctx = rf.Context(1 << 14, 1, 16, seed=1, engine=rf.RF_ENGINE_COMPAT, hist_mode=1)
    fh = ctx.flame(fl)
    h = ctx.histogram(64, 64)
    fh.warmup(1)
    with pytest.raises(rf.RefraktError, match="fast engine"):
        fh.draw_to_bins_noalloc(h, 1)
    ctx2 = rf.Context(1 << 14, 1, 16, seed=1, engine=rf.RF_ENGINE_FAST, hist_mode=1)
    f2 = ctx2.flame(fl)
    h2 = ctx2.histogram(64, 64)
    with pytest.raises(rf.RefraktError, match="warmup"):
        f2.draw_to_bins_noalloc(h2, 1)
