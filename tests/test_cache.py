"""Image and calibration cache formats (CPU): the files carry exactly what the
reference's readers check (openfwi/2_process.py:85-109), fingerprints are
SHA-256 of canonical JSON, and tampering invalidates a cache."""
import hashlib
import json

import numpy as np

from shotsweep_amd import cache


def test_fingerprint_is_sha256_of_canonical_json():
    value = {'b': [1, 2.5, 'x'], 'a': {'z': None, 'y': True}}
    text = '{"a":{"y":true,"z":null},"b":[1,2.5,"x"]}'
    assert cache.fingerprint(value) == hashlib.sha256(text.encode()).hexdigest()
    assert cache.fingerprint({'a': 1, 'b': 2}) == cache.fingerprint({'b': 2, 'a': 1})


def test_calibration_roundtrip_and_tamper(tmp_path):
    prov = {'schema': cache.CACHE_SCHEMA, 'rhos': [0.6, 1.0, 1.4]}
    meta = cache.write_calibration(tmp_path, np.sin(np.arange(1000.0)), np.linspace(0, 900, 70), prov)
    assert cache.calibration_valid(tmp_path, cache.fingerprint(prov))
    assert not cache.calibration_valid(tmp_path, cache.fingerprint({'other': 1}))
    with open(tmp_path / 'calibration_cache.json') as f:
        on_disk = json.load(f)
    assert set(on_disk) == {'fingerprint', 'provenance', 'wavelet_sha256', 'trend_sha256'} and on_disk == meta
    np.save(tmp_path / 'depth_trend.npy', np.zeros(70))          # tampered file
    assert not cache.calibration_valid(tmp_path, cache.fingerprint(prov))
    try:
        cache.write_calibration(tmp_path, np.zeros(1000), np.zeros(70), prov)
        raise AssertionError('zero wavelet accepted')
    except ValueError as e:
        assert 'invalid source wavelet' in str(e)


def test_family_cache_layout_and_validation(tmp_path):
    rng = np.random.default_rng(0)
    n = 5
    results = []
    for ks in ([0, 1], [2, 3, 4]):
        m = len(ks)
        results.append((ks, {
            'images': rng.standard_normal((m, 3, 5, 70, 70)).astype(np.float16),
            'surface_velocity': rng.uniform(1500, 2000, m).astype(np.float32),
            'surface_confidence': rng.uniform(0, 1, m).astype(np.float32),
            'surface_fit_rmse_ms': rng.uniform(0, 1, m).astype(np.float32),
            'surface_possible_refraction': rng.uniform(0, 1, m) > 0.5}))
    arrays = cache.family_arrays(n, results)
    calib = {'fingerprint': 'f' * 64, 'wavelet_sha256': 'a' * 64, 'trend_sha256': 'b' * 64}
    prov = cache.input_provenance(calib, 'FlatVel_A', {'gathers': 'c' * 64, 'models': 'd' * 64}, n, 3, 1)
    path = tmp_path / 'inputs' / 'FlatVel_A.npz'
    cache.write_family(path, arrays, prov)
    assert cache.inputs_valid(path, prov['fingerprint'], n)
    assert not cache.inputs_valid(path, prov['fingerprint'], n + 1)
    with np.load(path, allow_pickle=False) as c:
        assert set(c.files) == {'images', 'surface_velocity', 'surface_confidence', 'surface_fit_rmse_ms',
                                'surface_possible_refraction', 'cache_provenance'}
        assert c['images'].dtype == np.float16 and c['images'].shape == (n, 3, 5, 70, 70)
        assert c['surface_velocity'].dtype == np.float32 and c['surface_possible_refraction'].dtype == bool
        np.testing.assert_array_equal(c['images'][2:], results[1][1]['images'])
        meta = json.loads(str(c['cache_provenance'].item()))
    body = {k: v for k, v in meta.items() if k != 'fingerprint'}
    assert meta['fingerprint'] == cache.fingerprint(body) and meta['schema'] == 2
    bad = dict(arrays)
    bad['images'] = arrays['images'].copy()
    bad['images'][0, 0, 0, 0, 0] = np.float16(np.inf)
    cache.write_family(path, bad, prov)
    assert not cache.inputs_valid(path, prov['fingerprint'], n)
    assert not cache.inputs_valid(tmp_path / 'missing.npz', prov['fingerprint'], n)


def test_family_cache_passes_the_training_reader(tmp_path):
    """The consumer's provenance check (3_train.py:166-177), restated."""
    calib = {'fingerprint': 'f' * 64, 'wavelet_sha256': 'a' * 64, 'trend_sha256': 'b' * 64}
    man = {'n_total': 4, 'n_train': 2, 'n_val': 1}
    prov = cache.input_provenance(calib, 'CurveFault_A', {'gathers': 'c' * 64, 'models': 'd' * 64},
                                  man['n_total'], man['n_train'], man['n_val'])
    arrays = cache.family_arrays(4, [])
    path = tmp_path / 'CurveFault_A.npz'
    cache.write_family(path, arrays, prov)
    with np.load(path) as inp:
        metadata = json.loads(str(inp['cache_provenance'].item()))
        expected = dict(metadata)
        fp = expected.pop('fingerprint')
        assert (metadata['schema'] == 2 and metadata['family'] == 'CurveFault_A'
                and all(metadata[k] == man[k] for k in ('n_train', 'n_val', 'n_total'))
                and metadata['calibration_fingerprint'] == calib['fingerprint']
                and metadata['calibration_wavelet_sha256'] == calib['wavelet_sha256']
                and metadata['calibration_trend_sha256'] == calib['trend_sha256']
                and fp == hashlib.sha256(json.dumps(expected, sort_keys=True,
                                                    separators=(',', ':')).encode()).hexdigest())
        assert inp['images'].shape == (4, 3, 5, 70, 70) and inp['surface_velocity'].shape == (4,)


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
