"""On-disk formats of the imaging caches, compatible with the reference's readers.

The reference harness (openfwi/2_process.py:26-109, 137-218) writes two caches
that its training and test scripts consume (3_train.py:166-189):

* the calibration cache: ``source_wavelet.npy``, ``depth_trend.npy`` and
  ``calibration_cache.json`` {fingerprint, provenance, wavelet_sha256,
  trend_sha256};
* one image cache per family, ``inputs/<family>.npz`` with ``images`` float16
  (n, 3, 5, 70, 70), ``surface_velocity``, ``surface_confidence``,
  ``surface_fit_rmse_ms`` (float32, n), ``surface_possible_refraction`` (bool,
  n) and ``cache_provenance`` (a JSON string with its own fingerprint).

The fingerprint is the SHA-256 of the provenance serialised as canonical JSON
(sorted keys, no spaces); file digests are SHA-256 of the bytes.  Files are
written atomically (temporary file in the same directory, fsync, rename), so
a reader never sees a partial cache.  ``family_arrays`` assembles the cache
arrays from ``sweep.sweep_batch`` results.
"""
import hashlib
import json
import os
import tempfile
from pathlib import Path

import numpy as np

from .background import RHOS

CACHE_SCHEMA = 2


def file_digest(path):
    """SHA-256 hex digest of a file, read in 1 MiB blocks."""
    h = hashlib.sha256()
    with open(path, 'rb') as f:
        while True:
            block = f.read(1 << 20)
            if not block:
                break
            h.update(block)
    return h.hexdigest()


def fingerprint(value):
    """SHA-256 of canonical JSON (sorted keys, ',' and ':' separators)."""
    text = json.dumps(value, sort_keys=True, separators=(',', ':'))
    return hashlib.sha256(text.encode()).hexdigest()


def _atomic(path, write):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix=f'.{path.name}.', suffix='.tmp', dir=path.parent)
    try:
        with os.fdopen(fd, 'wb') as f:
            write(f)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)


def save_npy(path, values):
    _atomic(path, lambda f: np.save(f, values, allow_pickle=False))


def save_npz(path, **values):
    _atomic(path, lambda f: np.savez(f, **values))


def save_json(path, value):
    _atomic(path, lambda f: f.write(json.dumps(value, sort_keys=True, indent=2).encode()))


# ------------------------------------------------------------ calibration
def write_calibration(work, wavelet, trend, provenance):
    """Write the calibration cache; returns its metadata dict."""
    work = Path(work)
    wavelet, trend = np.asarray(wavelet), np.asarray(trend)
    if not np.isfinite(wavelet).all() or not np.any(wavelet) or not np.isfinite(trend).all():
        raise ValueError('calibration produced invalid source wavelet or depth trend')
    save_npy(work / 'source_wavelet.npy', wavelet)
    save_npy(work / 'depth_trend.npy', trend)
    meta = {'fingerprint': fingerprint(provenance), 'provenance': provenance,
            'wavelet_sha256': file_digest(work / 'source_wavelet.npy'),
            'trend_sha256': file_digest(work / 'depth_trend.npy')}
    save_json(work / 'calibration_cache.json', meta)
    return meta


def calibration_valid(work, expected_fingerprint):
    """True when the calibration cache matches the fingerprint and both files' digests."""
    work = Path(work)
    try:
        with open(work / 'calibration_cache.json') as f:
            meta = json.load(f)
        return (meta['fingerprint'] == expected_fingerprint
                and meta['wavelet_sha256'] == file_digest(work / 'source_wavelet.npy')
                and meta['trend_sha256'] == file_digest(work / 'depth_trend.npy'))
    except (OSError, ValueError, KeyError):
        return False


def calibrated(work, provenance, models, gathers, training_models):
    """Source wavelet and depth trend, from a valid calibration cache or computed and
    cached (2_process.py:137-161): the wavelet from ``gpu.calibrate`` on the
    calibration models and their gathers, the trend from the training models.
    Returns (wavelet, trend, metadata)."""
    from . import gpu
    from .background import depth_trend
    work = Path(work)
    if not calibration_valid(work, fingerprint(provenance)):
        wavelet = gpu.calibrate(models, gathers)
        trend = depth_trend(training_models)
        write_calibration(work, wavelet, trend, provenance)
    with open(work / 'calibration_cache.json') as f:
        meta = json.load(f)
    return (np.load(work / 'source_wavelet.npy', allow_pickle=False),
            np.load(work / 'depth_trend.npy', allow_pickle=False), meta)


# ------------------------------------------------------------ image cache
def family_arrays(n, results, nb=len(RHOS), ns=5, nz=70, nx=70):
    """Cache arrays for n models from (indices, sweep_batch result) pairs."""
    images = np.zeros((n, nb, ns, nz, nx), np.float16)
    surface = np.zeros(n, np.float32)
    confidence = np.zeros(n, np.float32)
    fit_rmse = np.zeros(n, np.float32)
    refraction = np.zeros(n, bool)
    for ks, r in results:
        ks = list(ks)
        images[ks] = r['images']
        surface[ks] = r['surface_velocity']
        confidence[ks] = r['surface_confidence']
        fit_rmse[ks] = r['surface_fit_rmse_ms']
        refraction[ks] = r['surface_possible_refraction']
    return {'images': images, 'surface_velocity': surface, 'surface_confidence': confidence,
            'surface_fit_rmse_ms': fit_rmse, 'surface_possible_refraction': refraction}


def input_provenance(calibration_meta, family, data_identity, n_total, n_train, n_val):
    """Provenance of a family cache, with its fingerprint (2_process.py:173-181)."""
    prov = {'schema': CACHE_SCHEMA, 'calibration_fingerprint': calibration_meta['fingerprint'],
            'calibration_wavelet_sha256': calibration_meta['wavelet_sha256'],
            'calibration_trend_sha256': calibration_meta['trend_sha256'],
            'family': family, 'data': data_identity, 'n_total': n_total, 'n_train': n_train, 'n_val': n_val}
    prov['fingerprint'] = fingerprint(prov)
    return prov


def write_family(path, arrays, provenance):
    """Write one family's image cache atomically."""
    save_npz(path, **arrays, cache_provenance=np.array(json.dumps(provenance, sort_keys=True)))


def inputs_valid(path, expected_fingerprint, n_total, nb=len(RHOS)):
    """True when a family cache has the fingerprint, shapes and finite values the
    reference's reader requires (2_process.py:97-109)."""
    try:
        with np.load(path, allow_pickle=False) as c:
            meta = json.loads(str(c['cache_provenance'].item()))
            return (meta['fingerprint'] == expected_fingerprint
                    and c['images'].shape == (n_total, nb, 5, 70, 70)
                    and c['surface_velocity'].shape == (n_total,)
                    and c['surface_confidence'].shape == (n_total,)
                    and np.isfinite(c['images']).all()
                    and np.isfinite(c['surface_velocity']).all()
                    and np.isfinite(c['surface_confidence']).all())
    except (OSError, ValueError, KeyError, EOFError):
        return False
