"""OpenFWI scores of velocity maps on the GPU (openfwi/physics/metrics.py:14-73).

MAE and RMSE of the maps normalised to [-1, 1] and SSIM of the same maps on
[0, 1] (11-pixel Gaussian window, sigma 1.5, zero padding, population moments,
data range 1), computed per map in float64 by ``ss_velocity_scores``.  The
window is formed on the host as the reference forms it (a centred arange,
exp(-x^2 / (2 * 1.5^2)), normalised, outer product).
"""
import numpy as np
import torch

from . import _lib
from .constants import VMAX, VMIN

WINDOW = 11


def window(size=WINDOW, sigma=1.5):
    """The 2-D SSIM window (metrics.py:39-42) as float64 (size, size)."""
    x = np.arange(size, dtype=np.float64) - size // 2
    g = np.exp(-x ** 2 / (2 * sigma ** 2))
    g = g / g.sum()
    return g[:, None] * g[None, :]


_WINDOWS = {}


def _window_on(device):
    w = _WINDOWS.get(device)
    if w is None:
        w = torch.tensor(window(), dtype=torch.float64, device=device).contiguous()
        _WINDOWS[device] = w
    return w


def velocity_scores(pred, true):
    """(B, H, W) device velocity maps (m/s, float32 or float64) -> (B, 3) float64
    tensor of (MAE, RMSE, SSIM) per map."""
    if pred.shape != true.shape or pred.dim() != 3 or pred.dtype != true.dtype:
        raise ValueError('scores expect matching (B, H, W) maps of one dtype')
    if not pred.is_cuda or not true.is_cuda:
        raise ValueError('scores take device tensors')
    if not (bool(torch.isfinite(pred).all()) and bool(torch.isfinite(true).all())):
        raise ValueError('Metrics require finite, matching 2-D velocity maps')
    pred, true = pred.contiguous(), true.contiguous()
    B, H, W = pred.shape
    out = torch.empty(B, 3, dtype=torch.float64, device=pred.device)
    _lib.check(_lib.fn('velocity_scores', pred.dtype)(B, H, W, _lib.ptr(pred), _lib.ptr(true), VMIN, VMAX,
                                                     _lib.ptr(_window_on(pred.device)), _lib.ptr(out),
                                                     _lib.stream_handle(pred.device)), 'velocity_scores')
    return out


def metrics(pred, true, device='cuda'):
    """Unrounded OpenFWI scores of two physical velocity maps, as a dict (metrics.py:65-73)."""
    a, b = np.asarray(pred, np.float64), np.asarray(true, np.float64)
    if a.ndim != 2 or a.shape != b.shape or not np.isfinite(a).all() or not np.isfinite(b).all():
        raise ValueError('Metrics require finite, matching 2-D velocity maps')
    s = velocity_scores(torch.tensor(a[None], device=device), torch.tensor(b[None], device=device))[0].cpu()
    return dict(mae=float(s[0]), rmse=float(s[1]), ssim=float(s[2]))
