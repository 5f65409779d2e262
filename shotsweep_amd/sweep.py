"""The per-model image sweep of the processing harness (openfwi/2_process.py:193-214).

For a batch of models with observed gathers: fit each model's surface
velocity (host), build len(RHOS) laterally constant backgrounds, compute the
per-shot RTM images of every (model, background, shot) on the GPU in one
batch (the observed gathers are uploaded once and shared by a model's
backgrounds), then post-process on the host exactly as the harness does:
zero the first two image rows, scale each model's images by the 99th
percentile of their absolute values, clip to +-4 and store as float16.

Caching, provenance fingerprints and atomic npz writes of the harness
(2_process.py:28-109, 216-218) are not part of this module.
"""
import numpy as np
import torch

from . import background as BG


def postprocess(img):
    """(n, K, ns, nz, nx) float images -> float16 cache entries (2_process.py:205-210).

    Raises ValueError on non-finite or zero-energy images, as the harness does."""
    img = np.array(img, dtype=np.float32, copy=True)
    n = img.shape[0]
    img[..., :2, :] = 0
    scale = np.percentile(np.abs(img).reshape(n, -1), 99, axis=1)[:, None, None, None, None]
    if not np.isfinite(img).all() or np.any(scale <= 0):
        raise ValueError('nonfinite or zero-energy images')
    return np.clip(img / scale, -4, 4).astype(np.float16)


def surface_estimates(obs_host, first_index=0):
    """Surface-velocity estimates per model; errors name the model (2_process.py:196-201)."""
    estimates = []
    for k, observed in enumerate(obs_host):
        try:
            estimates.append(BG.estimate_surface_velocity(observed))
        except ValueError as error:
            raise ValueError(f'model {first_index + k}: {error}') from error
    return estimates


def sweep_batch(sim, obs, trend, rhos=BG.RHOS, estimates=None, eps=1e-4, t_mute=0.12):
    """Images of a batch of models at len(rhos) backgrounds.

    obs: observed gathers (n, ns, nt, nx), NumPy (float64 as the harness reads
    them) or a device tensor.  Returns a dict with ``images`` (n, K, ns, nz, nx)
    float16 after post-processing, ``raw`` (the float32 images before it),
    ``surface_velocity``, ``surface_confidence``, ``surface_fit_rmse_ms`` and
    ``surface_possible_refraction``."""
    g = sim.geom
    obs_host = obs.detach().cpu().numpy() if torch.is_tensor(obs) else np.asarray(obs, np.float64)
    n = len(obs_host)
    if estimates is None:
        estimates = surface_estimates(obs_host)
    v1 = np.array([e['velocity'] for e in estimates])
    K = len(rhos)
    bgs = np.concatenate([BG.backgrounds(v, trend, rhos, nx=g.nx) for v in v1])
    prep = sim._prepare(bgs)
    observed = sim._observed(obs_host, n)                # uploaded once; row b uses model b // K
    vm = sim._tensor(np.repeat(v1, K)).contiguous()
    img, _ = sim.images_device(prep, observed, K, vm, t_mute=t_mute, eps=eps)
    raw = img.reshape(n, K, g.ns, g.nz, g.nx).cpu().numpy()
    return {
        'images': postprocess(raw),
        'raw': raw,
        'surface_velocity': v1.astype(np.float32),
        'surface_confidence': np.array([e['confidence'] for e in estimates], np.float32),
        'surface_fit_rmse_ms': np.array([e['fit_rmse_ms'] for e in estimates], np.float32),
        'surface_possible_refraction': np.array([e['possible_refraction'] for e in estimates], bool),
    }
