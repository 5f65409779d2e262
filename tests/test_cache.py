"""Image and calibration cache formats: the files carry exactly what the
reference's readers check (openfwi/2_process.py:85-109 for reuse,
3_train.py:166-177 for the training reader), fingerprints are SHA-256 of
canonical JSON, writes are atomic, and tampering invalidates a cache.  The GPU
test images a family in batches and writes its cache."""
import hashlib
import json

import numpy as np
import pytest

from shotsweep_amd import cache

CALIB = {'fingerprint': 'f' * 64, 'wavelet_sha256': 'a' * 64, 'trend_sha256': 'b' * 64}
DATA = {'gathers': 'c' * 64, 'models': 'd' * 64}


def test_fingerprint_is_sha256_of_canonical_json():
    value = {'b': [1, 2.5, 'x'], 'a': {'z': None, 'y': True}}
    text = '{"a":{"y":true,"z":null},"b":[1,2.5,"x"]}'
    assert cache.fingerprint(value) == hashlib.sha256(text.encode()).hexdigest()
    assert cache.fingerprint({'a': 1, 'b': 2}) == cache.fingerprint({'b': 2, 'a': 1})
    rec = cache.with_fingerprint({'b': 1, 'a': 2, 'fingerprint': 'stale'})
    assert rec['fingerprint'] == cache.fingerprint({'a': 2, 'b': 1})


def test_atomic_replace_keeps_the_old_file_on_failure(tmp_path):
    path = tmp_path / 'sub' / 'x.bin'
    with cache.replaced_atomically(path) as f:
        f.write(b'first')
    with pytest.raises(RuntimeError):
        with cache.replaced_atomically(path) as f:
            f.write(b'partial')
            raise RuntimeError('writer failed')
    assert path.read_bytes() == b'first' and sorted(p.name for p in path.parent.iterdir()) == ['x.bin']


def test_calibration_roundtrip_reuse_and_tamper(tmp_path):
    prov = {'schema': cache.CACHE_SCHEMA, 'rhos': [0.6, 1.0, 1.4]}
    cal = cache.CalibrationCache(tmp_path)
    assert cal.metadata() is None and not cal.matches(cache.fingerprint(prov))
    meta = cal.store(np.sin(np.arange(1000.0)), np.linspace(0, 900, 70), prov)
    assert cal.matches(cache.fingerprint(prov)) and not cal.matches(cache.fingerprint({'other': 1}))
    on_disk = json.loads((tmp_path / 'calibration_cache.json').read_text())
    assert set(on_disk) == {'fingerprint', 'provenance', 'wavelet_sha256', 'trend_sha256'} and on_disk == meta
    assert meta['wavelet_sha256'] == hashlib.sha256((tmp_path / 'source_wavelet.npy').read_bytes()).hexdigest()
    np.save(tmp_path / 'depth_trend.npy', np.zeros(70))          # tampered file
    assert not cal.matches(cache.fingerprint(prov))
    with pytest.raises(ValueError, match='invalid source wavelet'):
        cal.store(np.zeros(1000), np.zeros(70), prov)


def test_calibrated_reuses_a_valid_cache(tmp_path, monkeypatch):
    from shotsweep_amd import gpu
    calls = []
    monkeypatch.setattr(gpu, 'calibrate', lambda m, g: calls.append(1) or np.cos(np.arange(1000.0)))
    prov = {'schema': 2, 'calibration_indices': [0, 1]}
    models = np.full((2, 70, 70), 2000.0)
    w1, t1, m1 = cache.calibrated(tmp_path, prov, models, None, models)
    w2, t2, m2 = cache.calibrated(tmp_path, prov, models, None, models)
    assert len(calls) == 1 and np.array_equal(w1, w2) and np.array_equal(t1, t2) and m1 == m2
    cache.calibrated(tmp_path, {'schema': 2, 'calibration_indices': [0]}, models, None, models)
    assert len(calls) == 2


def _batch(rng, m):
    return {'images': rng.standard_normal((m, 3, 5, 70, 70)).astype(np.float16),
            'surface_velocity': rng.uniform(1500, 2000, m).astype(np.float32),
            'surface_confidence': rng.uniform(0, 1, m).astype(np.float32),
            'surface_fit_rmse_ms': rng.uniform(0, 1, m).astype(np.float32),
            'surface_possible_refraction': rng.uniform(0, 1, m) > 0.5}


def test_family_cache_layout_and_validation(tmp_path):
    rng = np.random.default_rng(0)
    n = 5
    fam = cache.FamilyCache(tmp_path / 'inputs' / 'FlatVel_A.npz', n)
    batches = [([0, 1], _batch(rng, 2)), ([2, 3, 4], _batch(rng, 3))]
    for ks, r in batches:
        fam.add(ks, r)
    prov = cache.family_provenance(CALIB, 'FlatVel_A', DATA, n, 3, 1)
    fam.write(prov)
    assert cache.FamilyCache.matches(fam.path, prov['fingerprint'], n)
    assert not cache.FamilyCache.matches(fam.path, prov['fingerprint'], n + 1)
    assert not cache.FamilyCache.matches(fam.path, 'e' * 64, n)
    with np.load(fam.path, allow_pickle=False) as c:
        assert set(c.files) == {'images', 'surface_velocity', 'surface_confidence', 'surface_fit_rmse_ms',
                                'surface_possible_refraction', 'cache_provenance'}
        assert c['images'].dtype == np.float16 and c['images'].shape == (n, 3, 5, 70, 70)
        assert c['surface_velocity'].dtype == np.float32 and c['surface_possible_refraction'].dtype == bool
        np.testing.assert_array_equal(c['images'][2:], batches[1][1]['images'])
        np.testing.assert_array_equal(c['surface_fit_rmse_ms'][:2], batches[0][1]['surface_fit_rmse_ms'])
        meta = json.loads(str(c['cache_provenance'].item()))
    body = {k: v for k, v in meta.items() if k != 'fingerprint'}
    assert meta == prov and meta['fingerprint'] == cache.fingerprint(body) and meta['schema'] == 2
    bad = cache.FamilyCache(fam.path, n)
    bad.add(range(n), _batch(rng, n))
    bad.arrays['images'][0, 0, 0, 0, 0] = np.float16(np.inf)
    bad.write(prov)
    assert not cache.FamilyCache.matches(fam.path, prov['fingerprint'], n)
    assert not cache.FamilyCache.matches(tmp_path / 'missing.npz', prov['fingerprint'], n)


def test_family_cache_passes_the_training_reader(tmp_path):
    """The consumer's provenance check (3_train.py:166-177), restated."""
    man = {'n_total': 4, 'n_train': 2, 'n_val': 1}
    prov = cache.family_provenance(CALIB, 'CurveFault_A', DATA, man['n_total'], man['n_train'], man['n_val'])
    fam = cache.FamilyCache(tmp_path / 'CurveFault_A.npz', 4)
    fam.write(prov)
    with np.load(fam.path) as inp:
        metadata = json.loads(str(inp['cache_provenance'].item()))
        expected = dict(metadata)
        fp = expected.pop('fingerprint')
        assert (metadata['schema'] == 2 and metadata['family'] == 'CurveFault_A'
                and all(metadata[k] == man[k] for k in ('n_train', 'n_val', 'n_total'))
                and metadata['calibration_fingerprint'] == CALIB['fingerprint']
                and metadata['calibration_wavelet_sha256'] == CALIB['wavelet_sha256']
                and metadata['calibration_trend_sha256'] == CALIB['trend_sha256']
                and fp == hashlib.sha256(json.dumps(expected, sort_keys=True,
                                                    separators=(',', ':')).encode()).hexdigest())
        assert inp['images'].shape == (4, 3, 5, 70, 70) and inp['surface_velocity'].shape == (4,)


@pytest.mark.gpu
def test_family_imaged_in_batches_equals_one_sweep(tmp_path):
    """Four models imaged in batches of 3 (GPU picks and host fits, then the batched
    sweep) and written through FamilyCache equal one sweep of all four with host
    picks, byte for byte in every array, and the written cache is reusable."""
    from shotsweep_amd import background as BG, gpu, sweep, synth
    import torch
    print('\n[case] family cache from batches of 3 vs one sweep', flush=True)
    models = np.stack([synth.flat_layers(seed=s) for s in (20, 21, 22, 23)])
    sim = gpu.Sim(device='cuda')
    obs = sim.gathers(models).astype(np.float64)
    trend = BG.depth_trend(models)
    fam = cache.FamilyCache(tmp_path / 'inputs' / 'FlatVel_A.npz', 4)
    for start in (0, 3):
        ks = list(range(start, min(start + 3, 4)))
        est = BG.estimate_surface_velocity_device(torch.as_tensor(obs[ks], device='cuda'))
        assert not any(isinstance(e, ValueError) for e in est)
        r = sweep.sweep_batch(sim, obs[ks], trend, estimates=est)
        fam.add(ks, r)
    prov = cache.family_provenance(CALIB, 'FlatVel_A', DATA, 4, 2, 1)
    fam.write(prov)
    whole = sweep.sweep_batch(sim, obs, trend)
    assert cache.FamilyCache.matches(fam.path, prov['fingerprint'], 4)
    with np.load(fam.path, allow_pickle=False) as c:
        for name, _, _ in cache.FAMILY_SCHEMA:
            np.testing.assert_array_equal(c[name], whole[name], err_msg=name)
