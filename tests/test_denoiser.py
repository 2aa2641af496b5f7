"""Denoiser network kernels and the U-Net against PyTorch fp32 references
(computed on the CPU from the same bf16-rounded inputs and weights)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPEC = [(16, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1), (128, 256, 1),
        (384, 128, 1), (192, 64, 1), (96, 32, 1), (32, 32, 1), (32, 16, 0)]


def _bf16(t):
    import torch
    return t.to(torch.bfloat16)


def _layers(p):
    """(weight [Cout, Cin, 3, 3], bias) per layer from the flat parameter vector."""
    import torch
    out, off = [], 0
    for cin, cout, _ in SPEC:
        w = torch.from_numpy(p[off:off + 9 * cin * cout].reshape(3, 3, cout, cin)).permute(2, 3, 0, 1)
        off += 9 * cin * cout
        b = torch.from_numpy(p[off:off + cout])
        off += cout
        out.append((w.contiguous(), b.clone()))
    assert off == p.size
    return out


def _unet_ref(p, x, round_bf16=True):
    """fp32 U-Net on NCHW x (CPU); activations rounded to bf16 like the GPU's."""
    import torch
    import torch.nn.functional as F
    Ls = _layers(p)
    r = (lambda t: t.to(torch.bfloat16).float()) if round_bf16 else (lambda t: t)
    wr = (lambda w: w.to(torch.bfloat16).float()) if round_bf16 else (lambda w: w)

    def conv(i, t):
        w, b = Ls[i]
        y = F.conv2d(t, wr(w), b, padding=1)
        return r(torch.relu(y) if SPEC[i][2] else y)

    pool = lambda t: r(F.avg_pool2d(t, 2))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    e0 = conv(1, conv(0, x))
    e1 = conv(3, conv(2, pool(e0)))
    e2 = conv(5, conv(4, pool(e1)))
    b = conv(6, pool(e2))
    d2 = conv(7, torch.cat([up(b), e2], 1))
    d1 = conv(8, torch.cat([up(d2), e1], 1))
    d0 = conv(10, conv(9, torch.cat([up(d1), e0], 1)))
    o = conv(11, d0)
    return x[:, :4] + o[:, :4]


@pytest.mark.parametrize("Cin,Cout,W,relu,cs_in,co_in,cs_out,co_out", [
    (16, 32, 70, 1, 16, 0, 32, 0),
    (32, 16, 64, 0, 48, 16, 24, 4),
    (96, 64, 40, 1, 96, 0, 96, 32),
    (64, 128, 130, 0, 64, 0, 128, 0),
])
def test_conv3x3_matches_torch(rf, gpu_available, Cin, Cout, W, relu, cs_in, co_in, cs_out, co_out):
    import torch
    import torch.nn.functional as F
    ctx = rf.Context(1 << 12, 1, 16)
    g = torch.Generator().manual_seed(Cin * 131 + Cout)
    N, H = 2, 9
    x = _bf16(torch.randn(N, H, W, cs_in, generator=g))
    w = _bf16(0.1 * torch.randn(3, 3, Cout, Cin, generator=g))
    b = torch.randn(Cout, generator=g)
    out = torch.zeros(N, H, W, cs_out, dtype=torch.bfloat16)
    xd, wd, bd, od = x.cuda(), w.cuda(), b.cuda(), out.cuda()
    ctx.check(ctx.L.rf_nn_conv3x3(ctx.h, xd.data_ptr(), cs_in, co_in, wd.data_ptr(), bd.data_ptr(), od.data_ptr(),
                                  cs_out, co_out, N, H, W, Cin, Cout, relu))
    got = od.cpu().float()
    ref = F.conv2d(x[..., co_in:co_in + Cin].float().permute(0, 3, 1, 2), w.float().permute(2, 3, 0, 1), b,
                   padding=1)
    if relu:
        ref = torch.relu(ref)
    ref = ref.permute(0, 2, 3, 1)
    g_view = got[..., co_out:co_out + Cout]
    err = (g_view - ref).abs()
    assert (err <= 2.0 ** -8 * ref.abs() + 1e-4).all(), float(err.max())
    rest = torch.cat([got[..., :co_out], got[..., co_out + Cout:]], -1)
    assert (rest == 0).all()  # channels outside the output view are untouched


# Channel tiles of the weight gradient: 32x16 (16, 32), 32x32 (96, 32),
# 64x32 (64, 128), 16x32 (32, 16), 16x16 (48, 16); runs of rows that cross
# 64-pixel strips (W 70, 130) and images; (384, 128, 16, 16): 32 rows, 8 runs.
@pytest.mark.parametrize("Cin,Cout,W,H", [(16, 32, 70, 10), (96, 32, 40, 10), (64, 128, 24, 10),
                                          (384, 128, 16, 16), (32, 16, 130, 6), (48, 16, 33, 3)])
def test_conv3x3_gradients_match_torch(rf, gpu_available, Cin, Cout, W, H):
    import torch
    import torch.nn.functional as F
    ctx = rf.Context(1 << 12, 1, 16)
    g = torch.Generator().manual_seed(Cin + 7 * Cout)
    N = 2
    x = _bf16(torch.randn(N, H, W, Cin, generator=g))
    dy = _bf16(torch.randn(N, H, W, Cout, generator=g))
    w = _bf16(0.1 * torch.randn(3, 3, Cout, Cin, generator=g))
    xr = x.float().permute(0, 3, 1, 2).clone().requires_grad_(True)
    wr = w.float().permute(2, 3, 0, 1).clone().requires_grad_(True)
    br = torch.zeros(Cout, requires_grad=True)
    F.conv2d(xr, wr, br, padding=1).backward(dy.float().permute(0, 3, 1, 2))
    # weight and bias gradients
    dw = torch.zeros(3, 3, Cout, Cin, device="cuda")
    db = torch.zeros(Cout, device="cuda")
    xd, dyd = x.cuda(), dy.cuda()
    ctx.check(ctx.L.rf_nn_conv3x3_wgrad(ctx.h, xd.data_ptr(), Cin, 0, dyd.data_ptr(), Cout, 0, dw.data_ptr(),
                                        db.data_ptr(), N, H, W, Cin, Cout))
    ref_dw = wr.grad.permute(2, 3, 0, 1)
    torch.testing.assert_close(dw.cpu(), ref_dw, rtol=1e-4, atol=1e-3)
    torch.testing.assert_close(db.cpu(), br.grad, rtol=1e-4, atol=1e-3)
    # input gradient: the same convolution kernel on flipped, transposed weights
    wt = torch.zeros(3, 3, Cin, Cout, dtype=torch.bfloat16, device="cuda")
    wd = w.cuda()
    ctx.check(ctx.L.rf_nn_flip_weights(ctx.h, wd.data_ptr(), wt.data_ptr(), Cin, Cout))
    assert torch.equal(wt.cpu(), w.flip(0, 1).permute(0, 1, 3, 2))
    dx = torch.zeros(N, H, W, Cin, dtype=torch.bfloat16, device="cuda")
    ctx.check(ctx.L.rf_nn_conv3x3(ctx.h, dyd.data_ptr(), Cout, 0, wt.data_ptr(), None, dx.data_ptr(), Cin, 0,
                                  N, H, W, Cout, Cin, 0))
    ref_dx = xr.grad.permute(0, 2, 3, 1)
    err = (dx.cpu().float() - ref_dx).abs()
    assert (err <= 2.0 ** -8 * ref_dx.abs() + 1e-4).all(), float(err.max())


@pytest.mark.parametrize("npix", [1, 257, 4096 + 3])
def test_log_density_input_matches_formula(rf, gpu_available, npix):
    """bf16(log10(1 + v scale)) in channels 0..3, the noise level
    bf16(log10(1 + scale)) in channel 4, zeros in 5..15; the
    one-pixel-per-thread kernel (16-byte aligned buffers) and the per-value
    kernel (an output 2 bytes off alignment) give the same bits."""
    import torch
    ctx = rf.Context(1 << 12, 1, 16)
    g = torch.Generator().manual_seed(npix)
    h = torch.rand(npix, 4, generator=g) * torch.tensor([40.0, 3.0, 0.5, 900.0])
    h[::7] = 0.0
    scale = 0.37
    hd = h.cuda()
    out_a = torch.full((npix * 16,), -1.0, dtype=torch.bfloat16, device="cuda")
    out_b = torch.full((npix * 16 + 1,), -1.0, dtype=torch.bfloat16, device="cuda")
    ctx.log_density_input(hd.data_ptr(), out_a.data_ptr(), npix, scale)
    ctx.log_density_input(hd.data_ptr(), out_b[1:].data_ptr(), npix, scale)
    a = out_a.cpu().view(npix, 16)
    b = out_b[1:].cpu().view(npix, 16)
    assert torch.equal(a.view(torch.int16), b.view(torch.int16))
    assert (a[:, 5:] == 0).all()
    nl = math.log10(1.0 + float(np.float32(scale)))
    assert (torch.abs(a[:, 4].double() - nl) <= 2.0 ** -8 * nl).all()
    ref = torch.log10(1.0 + h.double() * np.float32(scale))
    err = (a[:, :4].double() - ref).abs()
    assert (err <= 2.0 ** -8 * ref.abs() + 1e-6).all(), float(err.max())
    assert float(out_b[0].float()) == -1.0  # nothing written before the view


def test_adam_matches_formula(rf, gpu_available):
    import torch
    ctx = rf.Context(1 << 12, 1, 16)
    rng = np.random.default_rng(1)
    n = 1000
    p0 = rng.standard_normal(n).astype(np.float32)
    tp = torch.from_numpy(p0.copy()).requires_grad_(True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = (torch.from_numpy(p0.copy()).cuda(), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"))
    for step in range(1, 6):
        gr = rng.standard_normal(n).astype(np.float32)
        tp.grad = torch.from_numpy(gr)
        opt.step()
        gd = torch.from_numpy(gr).cuda()
        ctx.check(ctx.L.rf_nn_adam(ctx.h, p.data_ptr(), gd.data_ptr(), m.data_ptr(), v.data_ptr(), n, 1e-2, 0.9,
                                   0.999, 1e-8, step))
    torch.testing.assert_close(p.cpu(), tp.detach(), rtol=1e-5, atol=1e-6)


def _inputs(seed, N, H, W):
    import torch
    g = torch.Generator().manual_seed(seed)
    base = torch.rand(N, 4, H // 4, W // 4, generator=g) * 2.0
    base = torch.nn.functional.interpolate(base, scale_factor=4, mode="bilinear", align_corners=False)
    x = torch.zeros(N, 16, H, W)
    x[:, :4] = (base + 0.3 * torch.randn(N, 4, H, W, generator=g)).clamp(min=0)
    t = torch.zeros(N, 16, H, W)
    t[:, :4] = base
    return _bf16(x), _bf16(t)


def test_unet_forward_matches_torch(rf, gpu_available):
    import torch
    ctx = rf.Context(1 << 12, 1, 16)
    dn = ctx.denoiser(seed=3)
    assert 1.1e6 < dn.n_params < 1.3e6
    p = dn.params()
    N, H, W = 2, 32, 48
    x, _ = _inputs(5, N, H, W)
    xd = x.permute(0, 2, 3, 1).contiguous().cuda()
    y = torch.zeros(N, H, W, 4, device="cuda")
    dn.infer(xd.data_ptr(), y.data_ptr(), N, H, W)
    ref = _unet_ref(p, x.float()).permute(0, 2, 3, 1)
    got = y.cpu()
    rel = float((got - ref).norm() / ref.norm())
    corr = (got - x.float().permute(0, 2, 3, 1)[..., :4]).flatten()
    refc = (ref - x.float().permute(0, 2, 3, 1)[..., :4]).flatten()
    assert rel < 1e-2, rel
    assert float(torch.nn.functional.cosine_similarity(corr, refc, dim=0)) > 0.99


def test_unet_gradients_match_torch(rf, gpu_available):
    import torch
    ctx = rf.Context(1 << 12, 1, 16)
    dn = ctx.denoiser(seed=4)
    p = dn.params()
    p[-(9 * 32 * 16 + 16):] *= 10.0  # a larger output layer gives every layer a sizeable gradient
    dn.set_params(p)
    N, H, W = 1, 32, 32
    x, t = _inputs(6, N, H, W)
    xd = x.permute(0, 2, 3, 1).contiguous().cuda()
    td = t.permute(0, 2, 3, 1).contiguous().cuda()
    g, loss = dn.grads(xd.data_ptr(), td.data_ptr(), N, H, W)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    # reference: the same network with autograd through unrounded fp32
    import torch.nn.functional as F
    y = _unet_ref_autograd(pt, x.float())
    ref_loss = ((y - t.float()[:, :4]) ** 2).mean()
    ref_loss.backward()
    assert abs(loss - float(ref_loss)) <= 0.02 * float(ref_loss)
    gr = pt.grad.numpy()
    off = 0
    for i, (cin, cout, _) in enumerate(SPEC):
        n = 9 * cin * cout + cout
        a, b = g[off:off + n], gr[off:off + n]
        off += n
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
        assert cos > 0.99 and rel < 0.1, (i, cos, rel)


def _unet_ref_autograd(pt, x):
    import torch
    import torch.nn.functional as F
    Ls, off = [], 0
    for cin, cout, _ in SPEC:
        w = pt[off:off + 9 * cin * cout].reshape(3, 3, cout, cin).permute(2, 3, 0, 1)
        off += 9 * cin * cout
        b = pt[off:off + cout]
        off += cout
        Ls.append((w, b))

    def conv(i, t):
        y = F.conv2d(t, Ls[i][0], Ls[i][1], padding=1)
        return torch.relu(y) if SPEC[i][2] else y

    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    e0 = conv(1, conv(0, x))
    e1 = conv(3, conv(2, F.avg_pool2d(e0, 2)))
    e2 = conv(5, conv(4, F.avg_pool2d(e1, 2)))
    b = conv(6, F.avg_pool2d(e2, 2))
    d2 = conv(7, torch.cat([up(b), e2], 1))
    d1 = conv(8, torch.cat([up(d2), e1], 1))
    d0 = conv(10, conv(9, torch.cat([up(d1), e0], 1)))
    o = conv(11, d0)
    return x[:, :4] + o[:, :4]


def test_training_reduces_loss(rf, gpu_available):
    ctx = rf.Context(1 << 12, 1, 16)
    dn = ctx.denoiser(seed=9)
    N, H, W = 2, 64, 64
    x, t = _inputs(7, N, H, W)
    xd = x.permute(0, 2, 3, 1).contiguous().cuda()
    td = t.permute(0, 2, 3, 1).contiguous().cuda()
    losses = [dn.train_step(xd.data_ptr(), td.data_ptr(), N, H, W, 2e-3) for _ in range(60)]
    assert np.isfinite(losses).all()
    assert min(losses[-5:]) < 0.5 * losses[0], (losses[0], losses[-5:])


def test_trained_weights_denoise_held_out_flames(rf, gpu_available):
    """The committed weights bring a 4-spp log-density image of unseen flames
    closer to the same flames' 1024-spp image."""
    import torch
    from refrakt_amd import denoise
    ctx = rf.Context(1 << 12, 1, 16, seed=77)
    dn = ctx.denoiser()
    meta = denoise.load_weights(dn)
    assert meta["n_params"] == dn.n_params
    flames = denoise.training_flames(6, seed=99)  # a seed the training set never used
    S = 512
    X = torch.empty((6, S, S, 16), dtype=torch.bfloat16, device="cuda")
    T = torch.empty_like(X)
    denoise.render_log_density(ctx, flames, S, S, 4.0, X, 64)
    denoise.render_log_density(ctx, flames, S, S, 1024.0, T, 1024)
    y = torch.empty((1, S, S, 4), dtype=torch.float32, device="cuda")
    gain = []
    for i in range(6):
        dn.infer(X[i:i + 1].data_ptr(), y.data_ptr(), 1, S, S)
        t = T[i, :, :, :4].float()
        mse_in = float(((X[i, :, :, :4].float() - t) ** 2).mean())
        mse_out = float(((y[0] - t) ** 2).mean())
        gain.append(10 * np.log10(mse_in / mse_out))
    print("denoiser gain per flame (dB):", [round(g, 2) for g in gain])
    assert np.mean(gain) > 2.0 and min(gain) > 0.0
