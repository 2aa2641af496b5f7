"""Batched rendering (rf_batch_*): many frames in one persistent launch under a
captured HIP graph."""
import numpy as np
import pytest

from refrakt_amd import quality

pytestmark = pytest.mark.gpu


def _flame(rf, seed, n_xf, W, H):
    rng = np.random.default_rng(seed)
    xs = []
    for k in range(n_xf):
        th = rng.uniform(0, 2 * np.pi)
        s = rng.uniform(0.3, 0.8)
        aff = [s * np.cos(th), s * np.sin(th), -s * np.sin(th), s * np.cos(th)] + list(rng.uniform(-0.7, 0.7, 2))
        var = {"linear": 0.6, "spherical": 0.4} if k % 2 else {"linear": 0.5, "swirl": 0.2, "julia": 0.3}
        xs.append(rf.Xform(affine=aff, variations=var, weight=float(rng.uniform(0.2, 1)),
                           color=float(rng.uniform()), color_speed=0.5, opacity=1.0))
    pal = np.ones((256, 4), np.float32)
    pal[:, :3] = rng.uniform(0.1, 1, (256, 3))
    return rf.Flame(xforms=xs, palette=pal, size=[W, H], center=[0.0, 0.0], scale=min(W, H) / 4.0, rotate=0.0,
                    estimator_min=0, estimator_radius=5, estimator_curve=0.4, gamma=2.2, vibrancy=1.0,
                    brightness=4.0)


FRAMES = [(3, 256, 192, 2_000_000), (5, 320, 320, 3_000_000), (2, 200, 150, 1_000_000), (7, 128, 128, 500_000)]


def _frames(rf):
    return [(_flame(rf, 10 + i, n, W, H), W, H, S) for i, (n, W, H, S) in enumerate(FRAMES)]


def test_batch_counts_images_and_replays(rf, gpu_available):
    import torch
    ctx = rf.Context(1 << 12, 1, 16, seed=5)
    frames = _frames(rf)
    b = ctx.batch(frames)
    b.set_config(8, rf.RF_HIST_LDS, 64, 20)
    b.run()
    st = b.frame_stats()
    hists = [b.histogram(i) for i in range(len(frames))]
    for i, (fl, W, H, S) in enumerate(frames):
        items = -(-S // (512 * 64))
        d = int(st["drawn"][i])
        assert 0.5 * S <= d <= items * 512 * 64, (i, d)
        assert int(hists[i][..., 3].astype(np.float64).sum()) == d  # opacity 1: alpha counts samples
        # the graph's DE + tonemap equals the single-frame call on the same histogram
        h = ctx.histogram(W, H, device_ptr=b.hist_ptr(i))
        out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        ctx.density_tonemap(h, rf.DEParams(fl.estimator_radius, fl.estimator_min, fl.estimator_curve),
                            rf.TonemapParams(fl.gamma, fl.brightness, fl.vibrancy, 1e-4), out.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), b.image(i)), i
        h.close()
    assert (st["chaos_done_ms"] > 0).all() and (st["image_done_ms"] >= st["chaos_done_ms"]).all()
    assert (np.diff(st["image_done_ms"]) >= 0).all()  # DE launches run in frame order
    # a replay of the graph re-zeroes and recomputes: alpha counts are exact and equal
    b.run()
    st2 = b.frame_stats()
    assert np.array_equal(st["drawn"], st2["drawn"])
    for i in range(len(frames)):
        h2 = b.histogram(i)
        assert np.array_equal(h2[..., 3], hists[i][..., 3])
        np.testing.assert_allclose(h2[..., :3], hists[i][..., :3], rtol=2e-3, atol=1e-3)
    b.close()


def test_batch_lds_and_quad_agree(rf, gpu_available):
    ctx = rf.Context(1 << 12, 1, 16, seed=6)
    frames = _frames(rf)
    b = ctx.batch(frames)
    out = {}
    for mode in (rf.RF_HIST_QUAD, rf.RF_HIST_LDS):
        b.set_config(4, mode, 128, 20)
        b.run()
        out[mode] = ([b.histogram(i) for i in range(len(frames))], b.frame_stats()["drawn"].copy())
    hq, dq = out[rf.RF_HIST_QUAD]
    hl, dl = out[rf.RF_HIST_LDS]
    assert np.array_equal(dq, dl)
    for a, c in zip(hq, hl):
        assert np.array_equal(a[..., 3], c[..., 3])
        np.testing.assert_allclose(a[..., :3], c[..., :3], rtol=2e-3, atol=1e-3)
    b.close()


def test_batch_frame_matches_single_frame_statistically(rf, gpu_available):
    """A batched frame and the same flame through warmup/draw_to_bins sample the
    same attractor: their normalised log-density images agree."""
    fl = _flame(rf, 99, 4, 256, 256)
    S = 1 << 24
    ctx = rf.Context(1 << 16, 1, 16, seed=7)
    b = ctx.batch([(fl, 256, 256, S)])
    b.set_config(8, rf.RF_HIST_LDS, 256, 20)
    b.run()
    hb = b.histogram(0)
    db = int(b.frame_stats()["drawn"][0])
    fh = ctx.flame(fl)
    hist = ctx.histogram(256, 256)
    fh.warmup(19)
    ds = sum(fh.draw_to_bins(hist, 64) for _ in range(4))
    hs = hist.download()
    La = quality.log_density(hb, db, 4)
    Lb = quality.log_density(hs, ds, 4)
    p = quality.psnr(La, Lb)
    print(f"batched vs single-frame: PSNR {p:.2f} dB")
    assert p > 35.0
    b.close()


def test_batch_autotune_picks_a_candidate(rf, gpu_available):
    ctx = rf.Context(1 << 12, 1, 16, seed=8)
    b = ctx.batch(_frames(rf))
    cfg, ms = b.tune(0.5)
    assert cfg["block_waves"] in (4, 8) and cfg["hist_mode"] in (1, 4, 5) and cfg["iters"] in (64, 256, 1024)
    assert ms > 0
    b.run()
    assert b.frame_stats()["drawn"].min() > 0
    b.close()


@pytest.mark.parametrize("hist_mode", [1, 4])
def test_batch_program_matches_interpreter(rf, gpu_available, hist_mode):
    """Every frame's xform k has the same variations, so the batch runs one
    flame program for all frames; it computes the interpreter's bits: drawn
    counts and hit counts equal, colour sums within 1e-5 (atomic order)."""
    frames = _frames(rf)
    out = []
    for programs in (True, False):
        ctx = rf.Context(1 << 12, 1, 16, seed=5, flame_programs=programs)
        b = ctx.batch(frames)
        b.set_config(8, hist_mode, 64, 20)
        b.run()
        out.append((b.program_info()[0], b.frame_stats()["drawn"].copy(),
                    [b.histogram(i) for i in range(len(frames))]))
        b.close()
        ctx.close()
    (ua, da, ha), (ub, db, hb) = out
    assert ua == 1 and ub == 0
    assert np.array_equal(da, db)
    for a, c in zip(ha, hb):
        assert np.array_equal(a[..., 3], c[..., 3])
        np.testing.assert_allclose(a[..., :3], c[..., :3], rtol=1e-5, atol=1e-5)


def test_batch_with_mixed_structures_is_interpreted(rf, gpu_available):
    frames = _frames(rf)
    fl = frames[1][0]
    fl.xforms[0].variations = {"linear": 0.3, "heart": 0.7}  # xform 0 differs from the other frames'
    ctx = rf.Context(1 << 12, 1, 16, seed=5)
    b = ctx.batch(frames)
    b.set_config(8, rf.RF_HIST_LDS, 64, 20)
    b.run()
    assert b.program_info()[0] == 0
    assert (b.frame_stats()["drawn"] > 0).all()


def test_cfg4_workload_against_fp64_oracle(rf, gpu_available):
    """The cfg4 workload itself (64 animation frames of one flame family, 2-12
    xforms, 512^2 to 2048^2, 64-512 samples per pixel) in one graph replay:
    every frame's alpha sum equals its drawn count, and three frames (512^2 /
    12 xforms at 139 spp, 1024^2 / 10 xforms at 374 spp, 1920x1080 / 11
    xforms at 124 spp) score at least 40 dB against the FP64 oracle's
    log-density image at about 1024 spp (tests/golden/psnr_cfg4.npz; pool 2,
    pool 4 for the 1920-wide frame)."""
    import json
    import os

    import torch
    from refrakt_amd import configs
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    meta = json.load(open(os.path.join(golden, "psnr_cfg4.json")))
    refs = np.load(os.path.join(golden, "psnr_cfg4.npz"))
    frames = configs.cfg4()
    ctx = rf.Context(1 << 12, 1, 16, seed=12345)
    b = ctx.batch([(f.flame, f.width, f.height, f.samples) for f in frames])
    print("cfg4: 64 frames, one replay", flush=True)
    b.run()
    st = b.frame_stats()
    for i, f in enumerate(frames):
        h = b.histogram(i)
        d = int(st["drawn"][i])
        assert d >= 0.5 * f.samples, (i, d, f.samples)
        assert int(h[..., 3].astype(np.float64).sum()) == d, i
        fm = meta["frames"].get(str(i))
        if fm is not None:
            ref = quality.dequantize(refs[f"f{i}"], fm["peak_raw"])
            p = quality.psnr(quality.log_density(h, d, fm["pool"]), ref)
            print(f"cfg4 frame {i} ({f.width}x{f.height}, {len(f.flame.xforms)} xforms, "
                  f"{d / (f.width * f.height):.0f} spp): {p:.2f} dB vs FP64", flush=True)
            assert p >= 40.0, (i, p)
    assert sorted(int(k) for k in meta["frames"]) == [3, 10, 11]
    b.close()
    ctx.close()
    torch.cuda.synchronize()


def _sort_vs_lds(rf, b, frames):
    """Replays b in LDS mode, then in sort mode (capacities from that replay):
    the trajectories are the same, so every frame's hit counts agree exactly;
    colours within the sort path's fixed point (1 / (2 * 32767) per sample)
    plus the float adds of n partial sums into a bin (n 2^-24 |sum|)."""
    b.set_config(8, rf.RF_HIST_LDS, 64, 20)
    b.run()
    ref = [b.histogram(i) for i in range(len(frames))]
    d_ref = b.frame_stats()["drawn"].copy()
    b.set_config(8, rf.RF_HIST_SORT, 64, 20)
    b.run()
    d = b.frame_stats()["drawn"]
    assert np.array_equal(d, d_ref)
    for i in range(len(frames)):
        h = b.histogram(i)
        assert np.array_equal(h[..., 3], ref[i][..., 3]), i
        n = ref[i][..., 3:4]
        tol = 1e-6 + n / 65534.0 + n * 2.0 ** -24 * np.abs(ref[i][..., :3])
        assert (np.abs(h[..., :3] - ref[i][..., :3]) <= tol).all(), i
        assert int(h[..., 3].astype(np.float64).sum()) == int(d[i])


def test_batch_sort_path_matches_lds(rf, gpu_available):
    ctx = rf.Context(1 << 12, 1, 16, seed=5)
    frames = _frames(rf)
    b = ctx.batch(frames)
    print("batch sort vs LDS", flush=True)
    _sort_vs_lds(rf, b, frames)
    b.close()


def test_batch_sort_path_without_estimates(rf, gpu_available):
    """A batch whose first graph uses the sort path has no replay to size its
    record buffers from: uniform capacities overflow into quad atomics, and
    the counts still equal the LDS path's."""
    ctx = rf.Context(1 << 12, 1, 16, seed=5)
    frames = _frames(rf)
    b = ctx.batch(frames)
    b.set_config(8, rf.RF_HIST_SORT, 64, 20)   # captured before any replay
    b.run()
    got = [b.histogram(i)[..., 3].copy() for i in range(len(frames))]
    b.set_config(8, rf.RF_HIST_LDS, 64, 20)    # same batch: same random streams
    b.run()
    for i in range(len(frames)):
        assert np.array_equal(got[i], b.histogram(i)[..., 3]), i
    b.close()


def test_batch_slices_finish_frames_early(rf, gpu_available):
    """Frames of about 2^28 samples each: the replay runs in several slices
    (at most 2^29 samples), and the first frames' images are done well before
    the last's, because each slice's finished frames are density-estimated
    while the next slice draws.  Sort and LDS modes agree on every frame's
    counts across slices."""
    ctx = rf.Context(1 << 12, 1, 16, seed=9)
    frames = [(_flame(rf, 40 + i, 4, 512, 512), 512, 512, 260_000_000) for i in range(6)]
    b = ctx.batch(frames)
    print("batch slices", flush=True)
    _sort_vs_lds(rf, b, frames)
    st = b.frame_stats()
    lat = st["image_done_ms"]
    assert (np.diff(lat) >= 0).all()
    assert lat[0] < 0.6 * lat[-1], lat
    b.close()
