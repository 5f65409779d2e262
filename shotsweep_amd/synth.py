"""Seeded synthetic velocity models for the benchmark workloads.

No reference counterpart: the reference reads OpenFWI models from disk
(``openfwi/1_data.py``), which are not available offline.  These generators
produce models of the same kind at any grid size:

* ``flat_layers``  - 4 flat layers, 1500-4500 m/s increasing with depth
                     (single-model workload, 70 x 70);
* ``mixed``        - curved layers (2-6 interfaces, sinusoidal perturbation
                     amplitude <= 10 cells), one planar fault with random dip
                     and throw <= 15 cells, additive Gaussian-smoothed texture
                     (sigma 3 cells, std 5 % of local velocity), clipped to
                     1500-4500 m/s (batched workloads, any grid size);
* ``smooth``       - Gaussian smoothing (sigma 8 cells) of a model, the
                     starting model of the differentiation workload.
"""
import numpy as np
from scipy.ndimage import gaussian_filter

VMIN, VMAX = 1500.0, 4500.0


def flat_layers(seed=0, nz=70, nx=70, n_layers=4):
    """Flat layers with velocities increasing with depth."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(8, nz - 6), n_layers - 1, replace=False))
    speeds = np.sort(rng.uniform(VMIN, VMAX, n_layers))
    speeds[0], speeds[-1] = max(speeds[0], VMIN), min(speeds[-1], VMAX)
    z = np.arange(nz)
    layer = np.searchsorted(tops, z, side='right')
    return np.repeat(speeds[layer][:, None], nx, axis=1).astype(np.float64)


def _mixed_one(rng, nz, nx):
    n_if = int(rng.integers(2, 7))
    x = np.arange(nx)
    z = np.arange(nz)[:, None].astype(np.float64)
    base = np.sort(rng.uniform(4, nz - 4, n_if))
    amp = rng.uniform(0, 10, n_if)
    wavelength = rng.uniform(0.5, 2.0, n_if) * nx
    phase = rng.uniform(0, 2 * np.pi, n_if)
    depth = base[:, None] + amp[:, None] * np.sin(2 * np.pi * x[None] / wavelength[:, None] + phase[:, None])
    speeds = np.sort(rng.uniform(VMIN, VMAX, n_if + 1))

    def layered(zz):
        layer = (zz[None] >= depth[:, None, :]).sum(0)
        return speeds[layer]

    zz = np.broadcast_to(z, (nz, nx))
    model = layered(zz)
    # planar fault through a random point; cells on the hanging-wall side slip by `throw`
    z_f, x_f = rng.uniform(0.25, 0.75) * nz, rng.uniform(0.25, 0.75) * nx
    dip = rng.uniform(np.deg2rad(30), np.deg2rad(80)) * rng.choice((-1, 1))
    throw = rng.uniform(-15, 15)
    side = (np.arange(nx)[None] - x_f) * np.sin(dip) - (z - z_f) * np.cos(dip) > 0
    model = np.where(side, layered(zz - throw), model)
    # Gaussian-smoothed texture, 5 % of local velocity
    noise = gaussian_filter(rng.standard_normal((nz, nx)), 3.0, mode='reflect')
    noise /= noise.std() + 1e-12
    model = model * (1 + 0.05 * noise)
    return np.clip(model, VMIN, VMAX)


def mixed(n, seed=0, nz=70, nx=70):
    """``n`` seeded models of mixed structure, shape (n, nz, nx), float64."""
    rng = np.random.default_rng(seed)
    return np.stack([_mixed_one(rng, nz, nx) for _ in range(n)])


def smooth(models, sigma=8.0):
    """Gaussian-smoothed starting models (edge-reflecting), same shape."""
    models = np.asarray(models, np.float64)
    return np.stack([gaussian_filter(m, sigma, mode='nearest') for m in models])
