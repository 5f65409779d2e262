"""Sweep definition on the host: surface-velocity anchor, depth trend, backgrounds.

Caller-side logic of the path (openfwi/physics/background.py:10-99), in NumPy
on the host as in the reference.  ``estimate_surface_velocity`` fits
threshold onsets at 10-60 m offsets and rejects curved or inconsistent picks
with the reference's messages; ``backgrounds`` builds the three laterally
constant trial models surface velocity + rho * mean depth trend.
"""
import numpy as np

from .constants import DT, DX, NT, NX, SRC_X

RHOS = (0.6, 1.0, 1.4)


def _fit(offsets, times):
    """Least-squares line; returns (slope, RMS residual)."""
    slope, intercept = np.polyfit(offsets, times, 1)
    misfit = times - (intercept + slope * offsets)
    return slope, float(np.sqrt(np.mean(misfit ** 2)))


def _check(obs, frac):
    obs = np.asarray(obs)
    if obs.shape != (5, NT, NX) or not np.isfinite(obs).all():
        raise ValueError(f'observations must be finite with shape (5, {NT}, {NX})')
    if not 0 < frac < 1:
        raise ValueError('onset threshold frac must lie between zero and one')
    return obs


def first_breaks(obs, frac=0.05):
    """Per-trace peak |g| and first sample above frac * peak (background.py:32-40), on the host."""
    amplitude = np.abs(_check(obs, frac))
    peaks = amplitude.max(axis=1)                                   # (5, NX)
    return np.argmax(amplitude > frac * peaks[:, None, :], axis=1), peaks


def estimate_surface_velocity(obs, frac=0.05):
    """First-break surface velocity with quality diagnostics (background.py:19-83)."""
    onset, peaks = first_breaks(obs, frac)
    return estimate_from_picks(onset, peaks)


def estimate_surface_velocity_device(obs, frac=0.05):
    """Estimates for a batch of device gathers (B, 5, NT, NX): the picks run in a HIP
    kernel (ss_first_breaks), the fits on the host.  Returns one dict or ValueError
    per model, in order."""
    import torch
    from . import _lib
    if obs.dim() != 4 or tuple(obs.shape[1:]) != (5, NT, NX) or not bool(torch.isfinite(obs).all()):
        raise ValueError(f'observations must be finite with shape (5, {NT}, {NX})')
    if not 0 < frac < 1:
        raise ValueError('onset threshold frac must lie between zero and one')
    obs = obs.contiguous()
    B = obs.shape[0]
    onset = torch.empty(B, 5, NX, dtype=torch.int32, device=obs.device)
    peaks = torch.empty(B, 5, NX, dtype=obs.dtype, device=obs.device)
    _lib.check(_lib.fn('first_breaks', obs.dtype)(B * 5, NT, NX, _lib.ptr(obs), float(frac), _lib.ptr(onset),
                                                 _lib.ptr(peaks), _lib.stream_handle(obs.device)), 'first_breaks')
    onset, peaks = onset.cpu().numpy(), peaks.cpu().numpy().astype(np.float64)
    out = []
    for k in range(B):
        try:
            out.append(estimate_from_picks(onset[k], peaks[k]))
        except ValueError as error:
            out.append(error)
    return out


def estimate_from_picks(onset_index, peaks):
    """The fit and quality checks of background.py:41-83 from per-trace picks (5, NX)."""
    h = np.abs(np.arange(NX)[None, :] * DX - np.asarray(SRC_X)[:, None])
    onset = np.asarray(onset_index) * DT
    use = (h >= 10) & (h <= 250) & (peaks > 0)
    shots, recv = np.nonzero(use)
    offsets, arrivals = h[shots, recv], onset[shots, recv]
    if len(offsets) == 0:
        raise ValueError('surface velocity unavailable: no nonzero onset traces')
    distances = np.unique(offsets)
    times = np.array([np.median(arrivals[offsets == d]) for d in distances])
    near = (distances >= 10) & (distances <= 60)
    if near.sum() != 6:
        raise ValueError('surface velocity unavailable: insufficient 10–60 m offsets')
    slope, rmse = _fit(distances[near], times[near])
    if slope <= 0 or not 1400 <= 1 / slope <= 4600:
        raise ValueError('surface velocity unavailable: short-offset slope is outside 1400–4600 m/s')
    early = (distances >= 10) & (distances <= 40)
    late = (distances >= 30) & (distances <= 60)
    curvature = abs(_fit(distances[early], times[early])[0] - _fit(distances[late], times[late])[0]) / slope
    per_shot = []
    for s in range(5):
        keep = (shots == s) & (offsets >= 10) & (offsets <= 60)
        if len(np.unique(offsets[keep])) >= 4:
            per_shot.append(_fit(offsets[keep], arrivals[keep])[0])
    if len(per_shot) < 3:
        raise ValueError('surface velocity unavailable: fewer than three usable shots')
    spread = float(np.std(per_shot) / slope)
    if rmse > 0.002 or curvature > 0.15 or spread > 0.15:
        raise ValueError('surface velocity unreliable: curved or inconsistent short-offset arrivals '
                         f'(RMSE={rmse * 1000:.2f} ms, curvature={curvature:.3f}, shot spread={spread:.3f})')
    far = (distances >= 120) & (distances <= 250)
    far_slope = _fit(distances[far], times[far])[0] if far.sum() >= 4 else slope
    refraction = max(0.0, 1 - far_slope / slope)
    return {
        'velocity': float(np.clip(1 / slope, 1500.0, 4500.0)),
        'confidence': float(1 / (1 + rmse / DT + curvature / 0.15 + spread / 0.15 + refraction)),
        'fit_rmse_ms': rmse * 1000,
        'short_offset_curvature': float(curvature),
        'shot_slope_spread': spread,
        'far_offset_velocity': float(1 / far_slope if far_slope > 0 else float('nan')),
        'possible_refraction': bool(refraction > 0.15),
    }


def surface_velocity(obs, frac=0.05):
    """Velocity only; raises instead of clipping unusable fits (background.py:86-88)."""
    return estimate_surface_velocity(obs, frac)['velocity']


def depth_trend(models):
    """Mean over models of v(z) - v(0), lateral means first: (N, nz, nx) -> (nz,) (background.py:91-94)."""
    m = np.asarray(models, np.float64)
    return (m.mean(2) - m[:, :1, :].mean(2)).mean(0)


def backgrounds(v_surface, trend, rhos=RHOS, nx=NX):
    """Laterally constant slow/medium/fast backgrounds (K, nz, nx) (background.py:97-99)."""
    return np.stack([np.tile(np.clip(v_surface + r * trend, 1500.0, 4500.0)[:, None], (1, nx)) for r in rhos])
