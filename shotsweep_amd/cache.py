"""The cache files on either side of the imaging path, in the format the reference's
readers accept.

Format (shared with the reference because its training and test scripts read these
files; the code here is this project's own):

* calibration, in a work directory (written by 2_process.py:147-161):
  ``source_wavelet.npy`` and ``depth_trend.npy`` (float64), and
  ``calibration_cache.json`` = {fingerprint, provenance, wavelet_sha256, trend_sha256};
* one image cache per family, ``inputs/<family>.npz`` (2_process.py:187-218, read by
  3_train.py:166-189): the arrays of ``FAMILY_SCHEMA`` plus ``cache_provenance``, a
  JSON string (sorted keys) whose ``fingerprint`` is the fingerprint of its other keys;
* a fingerprint is the SHA-256 hex digest of canonical JSON (sorted keys, ',' and ':'
  separators); a file digest is the SHA-256 of its bytes.

Every file is replaced atomically (``replaced_atomically``), so a reader sees the
old file or the new one, never a partial write.  ``FamilyCache`` collects the
results of ``sweep.sweep_batch`` batch by batch and writes the family's file once
all models are in.
"""
import contextlib
import hashlib
import json
import os
import tempfile
from pathlib import Path

import numpy as np

from .background import RHOS

CACHE_SCHEMA = 2

# name, dtype and per-model shape of every array of a family cache (2_process.py:187-191, 216-218)
FAMILY_SCHEMA = (
    ('images', np.float16, (len(RHOS), 5, 70, 70)),
    ('surface_velocity', np.float32, ()),
    ('surface_confidence', np.float32, ()),
    ('surface_fit_rmse_ms', np.float32, ()),
    ('surface_possible_refraction', np.bool_, ()),
)
# arrays whose shape and finiteness the reference's validator checks (2_process.py:97-109)
CHECKED = ('images', 'surface_velocity', 'surface_confidence')


def sha256_of_file(path, block=1 << 20):
    """SHA-256 hex digest of a file's bytes."""
    h = hashlib.sha256()
    with open(path, 'rb') as f:
        for chunk in iter(lambda: f.read(block), b''):
            h.update(chunk)
    return h.hexdigest()


def fingerprint(value):
    """SHA-256 hex digest of ``value`` as canonical JSON."""
    return hashlib.sha256(json.dumps(value, sort_keys=True, separators=(',', ':')).encode()).hexdigest()


def with_fingerprint(record):
    """A copy of ``record`` with its own fingerprint (of every other key) added."""
    body = {k: v for k, v in record.items() if k != 'fingerprint'}
    return dict(body, fingerprint=fingerprint(body))


@contextlib.contextmanager
def replaced_atomically(path):
    """Yield a binary file whose contents replace ``path`` in one rename when the block
    exits cleanly (flushed and fsync'd first); if the block raises, ``path`` keeps its
    previous contents and the temporary file is removed."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    handle, temporary = tempfile.mkstemp(dir=path.parent, prefix=f'.{path.name}.', suffix='.tmp')
    done = False
    try:
        with os.fdopen(handle, 'wb') as f:
            yield f
            f.flush()
            os.fsync(f.fileno())
        os.replace(temporary, path)
        done = True
    finally:
        if not done and os.path.exists(temporary):
            os.remove(temporary)


# ------------------------------------------------------------ calibration
class CalibrationCache:
    """The calibrated source wavelet and depth trend of a work directory."""

    WAVELET, TREND, META = 'source_wavelet.npy', 'depth_trend.npy', 'calibration_cache.json'

    def __init__(self, work):
        self.work = Path(work)

    def metadata(self):
        """The parsed metadata file, or None when it is missing or unreadable."""
        try:
            return json.loads((self.work / self.META).read_text())
        except (OSError, ValueError):
            return None

    def matches(self, expected_fingerprint):
        """Whether the cache was written for this provenance fingerprint and both arrays
        still have the digests recorded with it."""
        meta = self.metadata()
        if not isinstance(meta, dict):
            return False
        try:
            return (meta['fingerprint'] == expected_fingerprint
                    and meta['wavelet_sha256'] == sha256_of_file(self.work / self.WAVELET)
                    and meta['trend_sha256'] == sha256_of_file(self.work / self.TREND))
        except (OSError, KeyError):
            return False

    def store(self, wavelet, trend, provenance):
        """Write both arrays, then the metadata that names their digests; returns it."""
        wavelet, trend = np.asarray(wavelet), np.asarray(trend)
        if not (np.isfinite(wavelet).all() and np.any(wavelet) and np.isfinite(trend).all()):
            raise ValueError('calibration produced invalid source wavelet or depth trend')
        for name, values in ((self.WAVELET, wavelet), (self.TREND, trend)):
            with replaced_atomically(self.work / name) as f:
                np.save(f, values, allow_pickle=False)
        meta = {'fingerprint': fingerprint(provenance), 'provenance': provenance,
                'wavelet_sha256': sha256_of_file(self.work / self.WAVELET),
                'trend_sha256': sha256_of_file(self.work / self.TREND)}
        with replaced_atomically(self.work / self.META) as f:
            f.write(json.dumps(meta, sort_keys=True, indent=2).encode())
        return meta

    def load(self):
        """(wavelet, trend, metadata) as stored."""
        return (np.load(self.work / self.WAVELET, allow_pickle=False),
                np.load(self.work / self.TREND, allow_pickle=False), self.metadata())

    def ensure(self, provenance, compute):
        """The cached calibration when it matches ``provenance``; otherwise
        ``compute()`` -> (wavelet, trend) is stored first."""
        if not self.matches(fingerprint(provenance)):
            self.store(*compute(), provenance)
        return self.load()


def calibrated(work, provenance, models, gathers, training_models):
    """Source wavelet and depth trend of a work directory (2_process.py:147-161): reused
    when the calibration cache matches ``provenance``, else the wavelet comes from
    ``gpu.calibrate`` on the calibration models and their gathers and the trend from
    the training models, and both are cached.  Returns (wavelet, trend, metadata)."""
    from . import gpu
    from .background import depth_trend
    return CalibrationCache(work).ensure(provenance,
                                         lambda: (gpu.calibrate(models, gathers), depth_trend(training_models)))


# ------------------------------------------------------------ image cache
def family_provenance(calibration, family, data_identity, n_total, n_train, n_val):
    """Provenance record of a family cache (its keys are the format's,
    2_process.py:176-183), with its fingerprint; ``calibration`` is the calibration
    cache's metadata."""
    return with_fingerprint({
        'schema': CACHE_SCHEMA, 'family': family, 'data': data_identity,
        'calibration_fingerprint': calibration['fingerprint'],
        'calibration_wavelet_sha256': calibration['wavelet_sha256'],
        'calibration_trend_sha256': calibration['trend_sha256'],
        'n_total': n_total, 'n_train': n_train, 'n_val': n_val})


class FamilyCache:
    """The image cache of one family of ``n_total`` models, filled batch by batch."""

    def __init__(self, path, n_total):
        self.path = Path(path)
        self.n_total = int(n_total)
        self.arrays = {name: np.zeros((self.n_total,) + shape, dtype) for name, dtype, shape in FAMILY_SCHEMA}

    def add(self, indices, result):
        """Store a ``sweep.sweep_batch`` result for the models ``indices``."""
        indices = list(indices)
        for name, dtype, _ in FAMILY_SCHEMA:
            self.arrays[name][indices] = np.asarray(result[name], dtype)

    def write(self, provenance):
        """Replace the family's file atomically with the arrays and the provenance."""
        text = json.dumps(provenance, sort_keys=True)
        with replaced_atomically(self.path) as f:
            np.savez(f, **self.arrays, cache_provenance=np.array(text))

    @staticmethod
    def matches(path, expected_fingerprint, n_total):
        """Whether the file at ``path`` is a readable family cache for this fingerprint
        with ``n_total`` models and finite values, the conditions under which the
        reference reuses a cache (2_process.py:97-109)."""
        shapes = {name: (n_total,) + shape for name, _, shape in FAMILY_SCHEMA}
        try:
            with np.load(path, allow_pickle=False) as c:
                if json.loads(str(c['cache_provenance'].item())).get('fingerprint') != expected_fingerprint:
                    return False
                return all(c[name].shape == shapes[name] and bool(np.isfinite(c[name]).all()) for name in CHECKED)
        except (OSError, ValueError, KeyError, EOFError):
            return False
