"""The processing harness's control flow (openfwi/2_process.py:67-82, 182-218): the
manifest checks, batching into one atomically written family cache, reuse of a
valid cache and the error that names the family and models.  The CPU tests stand
in for the GPU picks and sweep with functions of the model index; the GPU test
runs the whole harness on a small data root."""
import json

import numpy as np
import pytest

from shotsweep_amd import cache, process, sweep

N = 5


def _fake_sweep(sim, obs, trend, estimates=None):
    ids = obs[:, 0, 0, 0]                       # each fake gather carries its model index
    m = len(ids)
    images = np.broadcast_to(ids[:, None, None, None, None], (m, 3, 5, 70, 70)).astype(np.float16)
    return {'images': images, 'raw': images.astype(np.float32),
            'surface_velocity': (1500 + ids).astype(np.float32),
            'surface_confidence': np.full(m, 0.5, np.float32),
            'surface_fit_rmse_ms': np.full(m, 0.1, np.float32),
            'surface_possible_refraction': ids > 2}


def _gathers(n=N):
    g = np.zeros((n, 5, 4, 70))
    g[:, 0, 0, 0] = np.arange(n)
    return g


def _prov(n=N):
    calib = {'fingerprint': 'f' * 64, 'wavelet_sha256': 'a' * 64, 'trend_sha256': 'b' * 64}
    return cache.input_provenance(calib, 'FlatVel_A', {'gathers': 'c' * 64, 'models': 'd' * 64}, n, 2, 1)


@pytest.fixture
def fake_gpu(monkeypatch):
    calls = []
    monkeypatch.setattr(process, 'pick_estimates', lambda sim, obs, family, ks: [None] * len(ks))
    monkeypatch.setattr(sweep, 'sweep_batch', lambda *a, **k: calls.append(len(a[1])) or _fake_sweep(*a, **k))
    return calls


def test_family_batches_into_one_cache_and_reuses_it(tmp_path, fake_gpu):
    out, prov, log = tmp_path / 'FlatVel_A.npz', _prov(), []
    assert process.process_family(out, None, None, _gathers(), prov, batch=2, log=log.append)
    assert fake_gpu == [2, 2, 1] and log[-1].startswith('FlatVel_A: saved validated cache')
    with np.load(out, allow_pickle=False) as c:
        assert c['images'].shape == (N, 3, 5, 70, 70)
        np.testing.assert_array_equal(c['images'][:, 2, 4, 69, 69], np.arange(N))
        np.testing.assert_array_equal(c['surface_velocity'], 1500 + np.arange(N))
        np.testing.assert_array_equal(c['surface_possible_refraction'], np.arange(N) > 2)
        assert json.loads(str(c['cache_provenance'].item())) == prov
    assert not process.process_family(out, None, None, _gathers(), prov, batch=2, log=log.append)
    assert fake_gpu == [2, 2, 1] and log[-1] == 'FlatVel_A: matching validated cache already exists'


def test_failing_batch_keeps_the_previous_cache(tmp_path, fake_gpu, monkeypatch):
    out, prov = tmp_path / 'FlatVel_A.npz', _prov()
    process.process_family(out, None, None, _gathers(), prov, batch=2, log=lambda s: None)
    before = out.read_bytes()

    def failing(sim, obs, trend, estimates=None):
        if obs[0, 0, 0, 0] >= 2:
            raise ValueError('nonfinite or zero-energy images')
        return _fake_sweep(sim, obs, trend)
    monkeypatch.setattr(sweep, 'sweep_batch', failing)
    stale = _prov(N)
    stale['n_train'] = 3
    stale['fingerprint'] = cache.fingerprint({k: v for k, v in stale.items() if k != 'fingerprint'})
    with pytest.raises(ValueError, match=r'FlatVel_A models \[2, 3\]: nonfinite or zero-energy images; '
                                         r'no replacement cache saved'):
        process.process_family(out, None, None, _gathers(), stale, batch=2, log=lambda s: None)
    assert out.read_bytes() == before and list(tmp_path.iterdir()) == [out]


def test_pick_errors_name_the_model():
    obs = np.zeros((2, 5, 1000, 70))
    obs[1, 0, 0, 0] = np.nan
    with pytest.raises(ValueError, match=r'FlatVel_A model 8: observations must be finite'):
        process.pick_estimates(None, obs, 'FlatVel_A', [7, 8])


def _write_root(root, n_total=4, n_train=2, n_val=1, lengths=None):
    (root / 'data').mkdir(parents=True)
    fams = {}
    for f in ('FlatVel_A', 'CurveFault_A'):
        n = n_total if lengths is None else lengths[f]
        np.save(root / 'data' / f'{f}_gathers.npy', np.zeros((n, 5, 4, 70), np.float32))
        np.save(root / 'data' / f'{f}_models.npy', np.zeros((n, 1, 70, 70), np.float32))
        fams[f] = {'gathers': f'data/{f}_gathers.npy', 'models': f'data/{f}_models.npy'}
    man = {'n_total': n_total, 'n_train': n_train, 'n_val': n_val, 'families': fams}
    (root / 'data' / 'manifest.json').write_text(json.dumps(man))
    return man


def test_manifest_checks(tmp_path):
    man = _write_root(tmp_path / 'ok')
    m, g, v = process.load_manifest(tmp_path / 'ok')
    assert m == man and set(g) == {'FlatVel_A', 'CurveFault_A'} and isinstance(g['FlatVel_A'], np.memmap)
    _write_root(tmp_path / 'split', n_train=3, n_val=1)
    with pytest.raises(ValueError, match='nonempty train, validation, and test'):
        process.load_manifest(tmp_path / 'split')
    _write_root(tmp_path / 'short', lengths={'FlatVel_A': 4, 'CurveFault_A': 3})
    with pytest.raises(ValueError, match='CurveFault_A: data length does not match'):
        process.load_manifest(tmp_path / 'short')
    with pytest.raises(ValueError, match='unknown family'):
        process.process(tmp_path / 'ok', families='FlatVel_A,Style_B')


def test_calibration_provenance_tracks_data_and_code(tmp_path):
    man = _write_root(tmp_path)
    p1, data = process.calibration_provenance(tmp_path, man)
    assert set(p1['implementation']) >= {'process.py', 'gpu.py', 'background.py', 'sweep.py', 'libshotsweep.so'}
    assert p1['training_indices'] == [0, 1] and p1['calibration_indices'] == [0, 1] and p1['rhos'] == [0.6, 1.0, 1.4]
    assert data['FlatVel_A']['gathers'] == cache.file_digest(tmp_path / 'data' / 'FlatVel_A_gathers.npy')
    np.save(tmp_path / 'data' / 'FlatVel_A_gathers.npy', np.ones((4, 5, 4, 70), np.float32))
    p2, _ = process.calibration_provenance(tmp_path, man)
    assert cache.fingerprint(p1) != cache.fingerprint(p2)


@pytest.mark.gpu
def test_process_small_root_matches_one_batch_sweep(tmp_path):
    """Two families of 4 models: calibration, GPU picks and the sweep in batches of
    3 give the same cache bytes-for-bytes as one sweep of all 4 models with host
    picks; a second run reuses every cache."""
    from shotsweep_amd import background as BG, gpu, synth
    print('\n[case] process() on a 2-family, 4-model data root, batch 3', flush=True)
    root = tmp_path
    (root / 'data').mkdir()
    osim = gpu.Sim(device='cuda')
    fams = {}
    # models whose first breaks give reliable surface velocities (checked with the oracle's gathers)
    sets = {'FlatVel_A': np.stack([synth.flat_layers(seed=s) for s in (20, 21, 22, 23)]),
            'CurveFault_A': synth.mixed(4, seed=20)}
    for f, models in sets.items():
        models = models[:, None].astype(np.float32)
        np.save(root / 'data' / f'{f}_models.npy', models)
        np.save(root / 'data' / f'{f}_gathers.npy', osim.gathers(models[:, 0].astype(np.float64)).astype(np.float32))
        fams[f] = {'gathers': f'data/{f}_gathers.npy', 'models': f'data/{f}_models.npy'}
    (root / 'data' / 'manifest.json').write_text(json.dumps(
        {'n_total': 4, 'n_train': 2, 'n_val': 1, 'families': fams}))
    assert process.process(root, batch=3, log=print) == {'FlatVel_A': True, 'CurveFault_A': True}
    wavelet = np.load(root / 'work' / 'source_wavelet.npy')
    trend = np.load(root / 'work' / 'depth_trend.npy')
    sim = gpu.Sim(wavelet, device='cuda')
    for f in fams:
        obs = np.load(root / fams[f]['gathers']).astype(np.float64)
        whole = sweep.sweep_batch(sim, obs, trend)
        host_v = np.array([BG.estimate_surface_velocity(o)['velocity'] for o in obs], np.float32)
        with np.load(root / 'work' / 'inputs' / f'{f}.npz', allow_pickle=False) as c:
            np.testing.assert_array_equal(c['surface_velocity'], host_v)
            for key in ('images', 'surface_confidence', 'surface_fit_rmse_ms', 'surface_possible_refraction'):
                np.testing.assert_array_equal(c[key], whole[key])
    assert process.process(root, batch=3, log=print) == {'FlatVel_A': False, 'CurveFault_A': False}
