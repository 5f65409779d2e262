"""CPU restatement of the OpenFWI scores (openfwi/physics/metrics.py:14-73).

TEST INFRASTRUCTURE ONLY: the checker for the ss_velocity_scores kernels.
norm (metrics.py:14-16), torch_ssim (:19-52: 11-pixel Gaussian window,
sigma 1.5, zero padding, population moments, c1 = 0.01^2, c2 = 0.03^2) and
metrics (:65-73), in float64 with torch.nn.functional.conv2d.
Pinned by tests/test_oracle.py: a brute-force loop SSIM written from the
published formula, identity (SSIM 1, MAE = RMSE = 0) and constant offsets.
"""
import numpy as np
import torch
import torch.nn.functional as F

VMIN, VMAX = 1500.0, 4500.0


def norm(v):
    """metrics.py:14-16."""
    return (np.asarray(v, dtype=np.float64) - VMIN) / (VMAX - VMIN) * 2 - 1


def ssim(a, b, window_size=11):
    """SSIM of two (H, W) float64 maps on [0, 1] (metrics.py:19-52, size_average)."""
    a = torch.as_tensor(np.asarray(a, np.float64))[None, None]
    b = torch.as_tensor(np.asarray(b, np.float64))[None, None]
    x = torch.arange(window_size, dtype=torch.float64) - window_size // 2
    g = torch.exp(-x.square() / (2 * 1.5 ** 2))
    g = g / g.sum()
    w = (g[:, None] * g[None, :])[None, None]
    p = window_size // 2
    mu_a, mu_b = F.conv2d(a, w, padding=p), F.conv2d(b, w, padding=p)
    va = F.conv2d(a * a, w, padding=p) - mu_a.square()
    vb = F.conv2d(b * b, w, padding=p) - mu_b.square()
    cov = F.conv2d(a * b, w, padding=p) - mu_a * mu_b
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = ((2 * mu_a * mu_b + c1) * (2 * cov + c2)) / ((mu_a.square() + mu_b.square() + c1) * (va + vb + c2))
    return float(s.mean())


def metrics(pred, true):
    """metrics.py:65-73."""
    a, b = norm(pred), norm(true)
    d = a - b
    return dict(mae=float(np.abs(d).mean()), rmse=float(np.sqrt(np.square(d).mean())),
                ssim=ssim((a + 1) / 2, (b + 1) / 2))
