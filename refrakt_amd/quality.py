"""Image quality: the normalised log-density image and its PSNR.

A histogram (H x W RGBA, float) accumulated from `n_binned` samples is
density-estimated, scaled so that its mean over the image is independent of
the sample count, averaged over `pool` x `pool` blocks and mapped through
log10(1 + v) per channel:

    L = log10(1 + mean_pool(D) * W * H / n_binned)

PSNR against a reference image uses the reference's maximum as the peak:
10 log10(max(L_ref)^2 / mean((L - L_ref)^2)).  The functions take numpy
arrays or torch tensors (the GPU path computes on the device).
"""
from __future__ import annotations

import math


def pool_factor(width: int, max_width: int = 512) -> int:
    """Smallest power of two that brings the image width to max_width or less."""
    f = 1
    while width // f > max_width:
        f *= 2
    return f


def log_density(rgba, n_binned: int, pool: int, frame_pixels: int = 0):
    """Normalised log-density image, (H // pool, W // pool, 4).  For a window
    cut from a larger frame, frame_pixels is the frame's W * H (the
    normalisation is the frame's mean density)."""
    H, W = rgba.shape[0], rgba.shape[1]
    h, w = H // pool, W // pool
    x = rgba[:h * pool, :w * pool].reshape(h, pool, w, pool, 4)
    scale = float(frame_pixels or W * H) / float(n_binned)
    if hasattr(x, "double") and not hasattr(x, "astype"):  # torch
        import torch
        m = x.double().mean(dim=(1, 3)) * scale
        return torch.log10(1.0 + m)
    import numpy as np
    m = x.astype(np.float64).mean(axis=(1, 3)) * scale
    return np.log10(1.0 + m)


def image_to_bin_layout(img):
    """A density-estimated image (render-target layout) in the histogram's bin
    layout: texel (x, y) of the image holds bin (x + 1, H - 1 - y), and bin
    column 0 has no texel (its splat centre falls outside the view)."""
    if hasattr(img, "flip") and not hasattr(img, "astype"):
        import torch
        f = torch.flip(img, dims=(0,))
        out = torch.zeros_like(f)
        out[:, 1:] = f[:, :-1]
        return out
    import numpy as np
    f = img[::-1]
    out = np.zeros_like(f)
    out[:, 1:] = f[:, :-1]
    return out


def psnr(test, ref) -> float:
    """PSNR of `test` against `ref` in dB, peak = max(ref)."""
    d = test - ref
    if hasattr(d, "double") and not hasattr(d, "astype"):
        mse = float((d.double() ** 2).mean())
        peak = float(ref.max())
    else:
        import numpy as np
        mse = float(np.mean(np.square(d, dtype=np.float64)))
        peak = float(np.max(ref))
    if mse == 0.0:
        return float("inf")
    return 10.0 * math.log10(peak * peak / mse)


def quantize(L):
    """uint16 code of a non-negative image and the scale that decodes it."""
    import numpy as np
    L = np.asarray(L, np.float64)
    peak = float(L.max()) if L.size else 0.0
    q = np.zeros(L.shape, np.uint16) if peak == 0.0 else np.round(L / peak * 65535.0).astype(np.uint16)
    return q, peak


def dequantize(q, peak: float):
    import numpy as np
    return q.astype(np.float64) * (peak / 65535.0)
