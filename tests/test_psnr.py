"""Whole-image parity: the GPU's normalised log-density image against the FP64
oracle's, per benchmark workload.

The oracle images (tests/golden/psnr_<cfg>.npz, written by
tests/golden/make_psnr_refs.py) come from the FP64 chaos game at 2^32 binned
samples: the histogram itself (`raw`) and after double-precision density
estimation (`de`).  The GPU renders each workload's own job (cfg1 2^30, cfg2
2^32, cfg3 2^34 binned samples) with the fast engine, and its raw image is
scored against the oracle's raw image; where the sample counts agree (cfg2)
the density-estimated images are scored too (the estimator radii depend on
the absolute counts).  The runs use independent random streams, so the PSNR
measures the Monte Carlo noise of both plus any systematic difference; the
thresholds below are the stated bar.
"""
import json
import os

import numpy as np
import pytest

from refrakt_amd import quality

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# dB, normalised log-density image vs the FP64 oracle at 2^32 samples
PSNR_MIN_RAW = {"cfg1": 50.0, "cfg2": 50.0, "cfg3": 50.0}
PSNR_MIN_DE = {"cfg2": 50.0}


def _ref(name, kind="raw"):
    meta = json.load(open(os.path.join(GOLDEN, f"psnr_{name}.json")))
    q = np.load(os.path.join(GOLDEN, f"psnr_{name}.npz"))[kind]
    return meta, quality.dequantize(q, meta[f"peak_{kind}"])


@pytest.mark.parametrize("name", ["cfg1", "cfg2", "cfg3"])
def test_reference_images_are_consistent(name):
    for kind in ("raw", "de"):
        meta, ref = _ref(name, kind)
        pool = quality.pool_factor(meta["width"])
        assert meta["pool"] == pool
        assert ref.shape == (meta["height"] // pool, meta["width"] // pool, 4)
        assert abs(ref.max() - meta[f"peak_{kind}"]) < 1e-12 and ref.min() >= 0.0
    assert meta["attempted"] == 1 << 32 and 0.9 * meta["attempted"] <= meta["binned"] <= meta["attempted"]


def test_quality_functions_numpy_and_torch_agree():
    import torch
    rng = np.random.default_rng(3)
    a = rng.exponential(3.0, size=(64, 96, 4)).astype(np.float32)
    Ln = quality.log_density(a, 1000, 4)
    Lt = quality.log_density(torch.from_numpy(a), 1000, 4).numpy()
    assert Ln.shape == (16, 24, 4)
    np.testing.assert_allclose(Ln, Lt, rtol=1e-12, atol=1e-12)
    # closed form: a constant image of value v over n samples
    c = np.full((8, 8, 4), 2.0)
    np.testing.assert_allclose(quality.log_density(c, 64, 2), np.log10(1.0 + 2.0), rtol=1e-15)
    assert quality.psnr(Ln, Ln) == float("inf")
    noisy = Ln + 0.01 * Ln.max()
    assert abs(quality.psnr(noisy, Ln) - 40.0) < 1e-9
    q, peak = quality.quantize(Ln)
    assert np.abs(quality.dequantize(q, peak) - Ln).max() <= peak / 65535.0
    # bin layout of an image: texel (x, y) <- bin (x + 1, H - 1 - y)
    img = rng.uniform(size=(5, 7, 4))
    b = quality.image_to_bin_layout(img)
    assert np.array_equal(b[:, 0], np.zeros((5, 4)))
    assert np.array_equal(b[4 - 2, 3 + 1], img[2, 3])
    np.testing.assert_array_equal(quality.image_to_bin_layout(torch.from_numpy(img)).numpy(), b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cfg1", "cfg2", "cfg3"])
def test_log_density_psnr_vs_fp64_oracle(rf, gpu_available, name):
    import torch
    from refrakt_amd import configs
    meta, ref = _ref(name, "raw")
    wl = configs.WORKLOADS[name]()
    ctx = rf.Context(wl.particles, wl.temporal_samples, 1024, seed=4242, engine=rf.RF_ENGINE_FAST)
    fh = ctx.flame(wl.flame)
    hist = ctx.histogram(wl.width, wl.height)
    fh.warmup(wl.warmup_passes)
    drawn = sum(fh.draw_to_bins(hist, wl.draw_passes) for _ in range(wl.draw_calls))
    h = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    h.copy_(torch.from_numpy(hist.download()))
    L = quality.log_density(h, drawn, meta["pool"]).cpu().numpy()
    p = quality.psnr(L, ref)
    rmse = float(np.sqrt(np.mean((L - ref) ** 2))) / meta["peak_raw"]
    print(f"{name}: raw, GPU {drawn} binned samples vs FP64 oracle {meta['binned']}: "
          f"PSNR {p:.2f} dB, normalised RMSE {rmse:.3e}")
    assert p >= PSNR_MIN_RAW[name], (name, p)
    if name in PSNR_MIN_DE:
        meta, ref = _ref(name, "de")
        f = wl.flame
        ctx.density_estimate(hist, rf.DEParams(f.estimator_radius, f.estimator_min, f.estimator_curve),
                             h.data_ptr())
        L = quality.log_density(h, drawn, meta["pool"]).cpu().numpy()
        p = quality.psnr(L, ref)
        rmse = float(np.sqrt(np.mean((L - ref) ** 2))) / meta["peak_de"]
        print(f"{name}: density-estimated: PSNR {p:.2f} dB, normalised RMSE {rmse:.3e}")
        assert p >= PSNR_MIN_DE[name], (name, p)
