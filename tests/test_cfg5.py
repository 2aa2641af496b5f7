"""cfg5 on the GPU against the FP64 oracle: a held-out flame of the
denoiser's family at 2048x2048, samples per pixel doubling from 1 to 1024,
three methods scored on the frame's central 1024^2 window at pool 1 against
the oracle's image at 2^32 samples (tests/golden/psnr_cfg5_<seed>.npz):

  * raw samples: the chaos game's own histogram; at 256 spp it must be
    within 50 dB PSNR of the FP64 oracle (the parity bar for this config),
    and its curve must cross 40 dB inside the range (below at 1 spp);
  * plain (samples + density estimation, src/main.cpp:490-515): crosses
    40 dB inside the range (up to 1024 spp) and stays below the oracle's own
    DE floor (de_floor_db in the fixture's json) plus a noise margin;
  * denoised: more than 2 dB above the raw samples and above plain at 1-16
    spp, and reaches 40 dB before plain does.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cfg5_methods_against_fp64_oracle(rf, gpu_available):
    import torch
    from refrakt_amd import configs, convergence, denoise, quality
    seed = configs.CFG5_SEEDS[0]
    wl = configs.cfg5(seed)
    meta = json.load(open(os.path.join(GOLDEN, f"psnr_{wl.name}.json")))
    ref = quality.dequantize(np.load(os.path.join(GOLDEN, f"psnr_{wl.name}.npz"))["raw"], meta["peak_raw"])
    ref_t = torch.from_numpy(ref).cuda()
    ctx = rf.Context(wl.particles, 1, 16, seed=31, engine=rf.RF_ENGINE_FAST)
    dn = ctx.denoiser()
    denoise.load_weights(dn)
    print(f"cfg5 {wl.name}: curve 1..1024 spp", flush=True)
    res = convergence.quality_curve(ctx, wl, ref_t, meta["window"], [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024],
                                    dn)
    for p in res["curve"]:
        print(json.dumps(p))
    print(json.dumps(res["time_to_target"]))
    print(json.dumps(res["matched_wall"]))
    pts = {int(p["spp"]): p for p in res["curve"]}
    assert pts[256]["raw_db"] >= 50.0, pts[256]
    assert pts[1]["raw_db"] < 40.0 <= pts[256]["raw_db"]
    assert res["time_to_target"]["raw"]["bracketed"]
    assert res["time_to_target"]["plain"]["bracketed"]
    assert pts[1024]["plain_db"] <= meta["de_floor_db"] + 1.0
    for s in (1, 2, 4, 8, 16):
        assert pts[s]["denoised_db"] > pts[s]["raw_db"] + 2.0, pts[s]
        assert pts[s]["denoised_db"] > pts[s]["plain_db"], pts[s]
    tt = res["time_to_target"]
    assert tt["denoised"]["bracketed"] and tt["denoised"]["wall_s"] < tt["plain"]["wall_s"]
    dn.close()
    ctx.close()
