"""The processing harness's imaging run as library calls (openfwi/2_process.py:67-82, 110-218).

``process(root)`` reads ``root/data/manifest.json`` and each family's gathers
and models (memory-mapped .npy), calibrates the source wavelet and depth trend
from the training models only (or reuses a calibration cache whose fingerprint
and file digests match), then writes ``root/work/inputs/<family>.npz`` for
every family whose cache is missing or stale.  A family is imaged in batches of
models on the GPU: first-break picks in ``ss_first_breaks``, the fits on the
host, then the batched sweep of ``sweep.sweep_batch``.  Its cache is written
atomically once every model has passed, so a failure leaves the previous file
in place and the error names the family and the models.

The provenance holds SHA-256 digests of this implementation's modules and
library where the reference records its own sources, so a cache written by one
implementation is never taken for the other's.
"""
import json
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib, cache, gpu, sweep
from . import background as BG
from . import constants

DEFAULT_BATCH = 64           # models per sweep: 64 x 3 backgrounds x 5 shots = 960 propagations


def load_manifest(root):
    """Manifest, gathers and models of a data root (2_process.py:67-82); arrays are memory-mapped."""
    root = Path(root)
    with open(root / 'data' / 'manifest.json') as stream:
        manifest = json.load(stream)
    ntr, nval = manifest['n_train'], manifest.get('n_val', 0)
    if ntr < 1 or nval < 1 or ntr + nval >= manifest['n_total']:
        raise ValueError('manifest must define nonempty train, validation, and test partitions')
    gathers = {f: np.load(root / e['gathers'], mmap_mode='r', allow_pickle=False)
               for f, e in manifest['families'].items()}
    models = {f: np.load(root / e['models'], mmap_mode='r', allow_pickle=False)
              for f, e in manifest['families'].items()}
    for family in gathers:
        if len(gathers[family]) != manifest['n_total'] or len(models[family]) != manifest['n_total']:
            raise ValueError(f'{family}: data length does not match manifest n_total')
    return manifest, gathers, models


def implementation_digests():
    """SHA-256 of the modules and the kernel library that determine the images."""
    paths = [__file__, gpu.__file__, BG.__file__, constants.__file__, sweep.__file__, _lib.LIB_PATH]
    return {Path(p).name: cache.file_digest(p) for p in paths}


def calibration_provenance(root, manifest):
    """(calibration provenance, per-family data digests) of a data root (2_process.py:131-144)."""
    root = Path(root)
    ntr = manifest['n_train']
    data = {f: {key: cache.file_digest(root / e[key]) for key in ('gathers', 'models')}
            for f, e in manifest['families'].items()}
    provenance = {'schema': cache.CACHE_SCHEMA, 'implementation': implementation_digests(),
                  'manifest': manifest, 'data': data, 'training_indices': list(range(ntr)),
                  'calibration_indices': list(range(min(2, ntr))), 'rhos': list(BG.RHOS)}
    return provenance, data


def pick_estimates(sim, obs, family, ks):
    """Surface-velocity estimates of a batch of observed gathers (n, 5, NT, NX) float64.

    The picks run on the GPU in float64; a model whose gathers or fit fail raises
    ValueError naming the family and the model (2_process.py:196-201)."""
    for k, observed in zip(ks, obs):
        try:
            BG._check(observed, 0.05)
        except ValueError as error:
            raise ValueError(f'{family} model {k}: {error}; no replacement family cache saved') from error
    estimates = BG.estimate_surface_velocity_device(torch.as_tensor(obs, dtype=torch.float64, device=sim.device))
    for k, estimate in zip(ks, estimates):
        if isinstance(estimate, ValueError):
            raise ValueError(f'{family} model {k}: {estimate}; no replacement family cache saved') from estimate
    return estimates


def process_family(out, sim, trend, gathers, provenance, batch=DEFAULT_BATCH, log=print):
    """Write one family's image cache unless a valid one exists (2_process.py:182-218).

    gathers: the family's (n_total, 5, NT, NX) array (memory-mapped is fine).
    Returns True when the cache was written, False when a valid one was reused."""
    family, n = provenance['family'], provenance['n_total']
    if batch < 1:
        raise ValueError('batch must be positive')
    if cache.inputs_valid(out, provenance['fingerprint'], n):
        log(f'{family}: matching validated cache already exists')
        return False
    results = []
    started = time.time()
    for i in range(0, n, batch):
        ks = list(range(i, min(i + batch, n)))
        obs = np.asarray(gathers[ks[0]:ks[-1] + 1], np.float64)
        estimates = pick_estimates(sim, obs, family, ks)
        try:
            r = sweep.sweep_batch(sim, obs, trend, estimates=estimates)
        except ValueError as error:
            raise ValueError(f'{family} models {ks}: {error}; no replacement cache saved') from error
        r.pop('raw', None)
        results.append((ks, r))
        log(f'{family}: {ks[-1] + 1}/{n} models, {time.time() - started:.0f}s')
    cache.write_family(out, cache.family_arrays(n, results), provenance)
    log(f'{family}: saved validated cache in {time.time() - started:.0f}s')
    return True


def process(root, families='all', batch=DEFAULT_BATCH, device='cuda', log=print):
    """Calibrate and image every requested family of a data root (2_process.py:110-218).

    families: 'all', a comma-separated string or a list of names.  Returns
    {family: True if its cache was written, False if a valid one was reused}."""
    root = Path(root).resolve()
    work = root / 'work'
    manifest, gathers, models = load_manifest(root)
    if families == 'all':
        names = list(gathers)
    else:
        names = families.split(',') if isinstance(families, str) else list(families)
    if not set(names) <= set(gathers):
        raise ValueError('unknown family requested')
    ntr = manifest['n_train']
    (work / 'inputs').mkdir(parents=True, exist_ok=True)
    provenance, data = calibration_provenance(root, manifest)
    idx = range(min(2, ntr))
    wavelet, trend, meta = cache.calibrated(
        work, provenance,
        [np.array(models[f][k, 0], float) for f in gathers for k in idx],
        [np.array(gathers[f][k], float) for f in gathers for k in idx],
        np.concatenate([np.asarray(models[f][:ntr, 0], np.float32) for f in gathers]))
    sim = gpu.Sim(wavelet, device=device)
    written = {}
    for family in names:
        prov = cache.input_provenance(meta, family, data[family], manifest['n_total'], ntr, manifest['n_val'])
        written[family] = process_family(work / 'inputs' / f'{family}.npz', sim, trend, gathers[family],
                                         prov, batch, log)
    return written
