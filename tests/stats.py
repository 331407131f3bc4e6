"""Statistical image comparison used by the parity tests.

Per-tile z-test: the difference of two independent Monte Carlo tile means
divided by the standard error estimated from the per-pixel second moments.
The standard error has a floor for float32 accumulation: both renderers sum
a pixel's samples in float32 (as the reference's frame buffer does,
sample_renderer.cpp:104), and n additions of one constant round by up to
about n * 2^-24 relative, always in the same direction.  Tiles whose samples
are all equal (a constant sky seen directly) have zero variance, so their
means may differ by that rounding alone; the floor is (n_a + n_b) * 2^-24 of
the tile mean, and at least 2e-6.
relMSE over tile means: mean((a - b)^2 / (b^2 + eps)).
"""
import numpy as np


def tile_view(img, t):
    H, W = img.shape[:2]
    H2, W2 = H - H % t, W - W % t
    return img[:H2, :W2].reshape(H2 // t, t, W2 // t, t, -1)


def tile_mean_and_var(mean, sumsq, spp, t):
    """Tile means and the variance of those means."""
    var_px = np.maximum(sumsq / spp - mean ** 2, 0.0) / spp  # variance of each pixel mean
    m = tile_view(mean, t).mean(axis=(1, 3))
    v = tile_view(var_px, t).sum(axis=(1, 3)) / (t * t) ** 2
    return m, v


def compare(a_mean, a_sq, a_spp, b_mean, b_sq, b_spp, tile=8, eps=1e-4):
    ma, va = tile_mean_and_var(a_mean.astype(np.float64), a_sq.astype(np.float64), a_spp, tile)
    mb, vb = tile_mean_and_var(b_mean.astype(np.float64), b_sq.astype(np.float64), b_spp, tile)
    rounding = max(2e-6, (a_spp + b_spp) * 2.0 ** -24)
    floor = (rounding * np.maximum(np.abs(ma), np.abs(mb))) ** 2
    z = (ma - mb) / np.sqrt(va + vb + floor + 1e-20)
    rel = ((ma - mb) ** 2 / (mb ** 2 + eps)).mean()
    return z, rel
