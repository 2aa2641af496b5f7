"""rf_render, the drop-in render(flame, params) (src/main.cpp:403-535): zero the
histogram, warmup, draw_calls x draw_to_bins, density estimation, tonemap.

Checked through the C ABI against the same steps called one by one on a
second context with the same seed and modes:
  * alpha (hit counts) bit-exact; colour sums within 2 n 2^-24 |sum| for a
    bin of n samples (float atomics add in any order);
  * stats.drawn = the steps' drawn counts = the alpha sum (opacity 1);
  * stats.iterations = P * (warmup_passes + 1 + draw_passes * draw_calls);
  * the image equals rf_density_tonemap of rf_render's own histogram bit for
    bit (the render's separate DE and tonemap kernels compute the fused
    kernel's values);
  * both device timings are reported.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flame(rf, W, H, seed=31):
    rng = np.random.default_rng(seed)
    xs = []
    for k in range(4):
        th = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(0.4, 0.8)
        xs.append(rf.Xform(affine=[s * math.cos(th), s * math.sin(th), -s * math.sin(th), s * math.cos(th)]
                           + list(rng.uniform(-0.7, 0.7, 2)),
                           variations={"linear": 0.6, "spherical": 0.3, "swirl": 0.1} if k % 2 else {"linear": 1.0},
                           weight=float(rng.uniform(0.3, 1.0)), color=float(rng.uniform()), color_speed=0.5,
                           opacity=1.0))
    fin = rf.Xform(affine=[0.95, 0.05, -0.05, 0.95, 0.0, 0.0], variations={"linear": 1.0}, color=0.5,
                   color_speed=0.1, opacity=1.0)
    pal = np.ones((256, 4), np.float32)
    pal[:, 0] = np.linspace(0, 1, 256)
    pal[:, 1] = np.abs(np.sin(np.linspace(0, 3, 256)))
    pal[:, 2] = np.linspace(1, 0, 256)
    return rf.Flame(xforms=xs, final_xform=fin, palette=pal, size=[W, H], center=[0.0, 0.0], scale=W / 4.5,
                    estimator_radius=5, estimator_min=0, estimator_curve=0.4, gamma=2.2, vibrancy=1.0,
                    brightness=4.0)


@pytest.mark.parametrize("engine,hist_mode,T", [("fast", 1, 1), ("fast", 0, 1), ("compat", 1, 4)])
@pytest.mark.parametrize("de_on", [True, False])
def test_render_equals_its_steps(rf, gpu_available, engine, hist_mode, T, de_on):
    import torch
    P, W, H = 1 << 16, 320, 256
    warm, passes, calls = 7, 6, 3
    eng = rf.RF_ENGINE_FAST if engine == "fast" else rf.RF_ENGINE_COMPAT
    fl = _flame(rf, W, H)
    print(f"render: engine={engine} hist_mode={hist_mode} T={T} de={de_on}", flush=True)
    ctx_a = rf.Context(P, T, 16, seed=5, engine=eng, hist_mode=hist_mode, block_waves=8)
    hist, img, st = rf.render(ctx_a, fl, W, H, warmup_passes=warm, draw_passes=passes, draw_calls=calls,
                              density_estimation=de_on)

    ctx_b = rf.Context(P, T, 16, seed=5, engine=eng, hist_mode=hist_mode, block_waves=8)
    fh = ctx_b.flame(fl)
    h = ctx_b.histogram(W, H)
    fh.warmup(warm, tss_width=1.2 / 60.0)
    drawn = sum(fh.draw_to_bins(h, passes) for _ in range(calls))
    g = h.download()

    assert st.drawn == drawn > 0
    assert int(hist[..., 3].sum(dtype=np.float64)) == drawn
    assert np.array_equal(hist[..., 3], g[..., 3]), "hit counts differ between rf_render and its steps"
    # two float-atomic sums of a bin's n samples: |difference| <= 2 n 2^-24 |sum|
    tol = 1e-6 + 2.0 * g[..., 3:4] * 2.0 ** -24 * np.abs(g[..., :3])
    assert (np.abs(hist[..., :3] - g[..., :3]) <= tol).all()
    assert st.iterations == P * (warm + 1 + passes * calls)
    assert st.ms_warmup > 0 and st.ms_draw > 0 and st.ms_density > 0 and st.ms_tonemap > 0

    h2 = ctx_b.histogram(W, H)
    h2.upload(hist)
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    de = rf.DEParams(fl.estimator_radius, fl.estimator_min, fl.estimator_curve) if de_on else rf.DEParams(0, 0, 0.0)
    tm = rf.TonemapParams(fl.gamma, fl.brightness, fl.vibrancy, 1e-4)
    ctx_b.density_tonemap(h2, de, tm, out.data_ptr())
    torch.cuda.synchronize()
    ref = out.cpu().numpy()
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), "image differs from rf_density_tonemap"
    assert img[..., :3].max() > 0.1
    for o in (h2, h, fh):
        o.close()
    ctx_b.close()
    ctx_a.close()
