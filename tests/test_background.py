"""Host-side sweep definition (shotsweep_amd.background) against the reference's
own surface-velocity checks (tests/test_physics.py:91-125 of the reference),
with gathers from the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import fd
from shotsweep_amd import background as BG


@pytest.fixture(scope='module')
def observed():
    torch.set_num_threads(4)
    models = [np.full((70, 70), v, dtype=float) for v in (1500, 2500, 4500)]
    for depth in (5, 2):
        m = np.full((70, 70), 1500.0)
        m[depth:] = 4500.0
        models.append(m)
    return fd.OracleSim().gathers(np.array(models), chunk=5)


def test_homogeneous_velocities_and_quality(observed):
    for index, truth in enumerate((1500, 2500, 4500)):
        est = BG.estimate_surface_velocity(observed[index])
        assert abs(est['velocity'] - truth) / truth < 0.03
        assert 0 < est['confidence'] <= 1
        assert not est['possible_refraction']


def test_shallow_refractor_uses_short_offsets(observed):
    est = BG.estimate_surface_velocity(observed[3])
    assert abs(est['velocity'] - 1500) < 100
    assert est['possible_refraction'] and est['confidence'] < 0.6


def test_unresolved_shallow_refractor_rejected(observed):
    with pytest.raises(ValueError, match='unreliable'):
        BG.estimate_surface_velocity(observed[4])


def test_zero_and_nonfinite_inputs_rejected(observed):
    with pytest.raises(ValueError, match='no nonzero'):
        BG.estimate_surface_velocity(np.zeros((5, 1000, 70)))
    bad = observed[0].copy()
    bad[0, 0, 0] = np.nan
    with pytest.raises(ValueError, match='finite'):
        BG.estimate_surface_velocity(bad)


def test_depth_trend_and_backgrounds_closed_form():
    models = np.stack([np.tile(np.linspace(1500, 3000, 70)[:, None], (1, 70)) + k * 100 for k in range(3)])
    trend = BG.depth_trend(models)
    np.testing.assert_allclose(trend, np.linspace(0, 1500, 70), atol=1e-9)
    bgs = BG.backgrounds(2000.0, trend)
    assert bgs.shape == (3, 70, 70)
    np.testing.assert_allclose(bgs[1, :, 7], np.clip(2000 + trend, 1500, 4500))
    np.testing.assert_allclose(bgs[2, -1, 0], min(2000 + 1.4 * 1500, 4500))
    assert np.all(bgs == bgs[:, :, :1])          # laterally constant


# --------------------------------------------- values pinned to the reference
NAMES = ['homogeneous_1500', 'homogeneous_2500', 'homogeneous_4500', 'refractor_depth5', 'refractor_depth2']


def _surface_golden():
    import json
    from pathlib import Path
    path = Path(__file__).resolve().parent / 'golden' / 'reference_surface_velocity.json'
    return json.loads(path.read_text())


def _digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _same_estimate(got, want):
    """The dict of the reference's estimate_surface_velocity, every value equal
    (exact float equality: the same picks, medians and fits), or its error."""
    if 'error' in want:
        assert isinstance(got, ValueError) and str(got) == want['error'], (got, want)
        return
    assert not isinstance(got, Exception), got
    assert set(got) == set(want)
    for k, v in want.items():
        assert got[k] == v or (np.isnan(v) and np.isnan(got[k])), (k, got[k], v)


def _host_estimate(obs):
    try:
        return BG.estimate_surface_velocity(obs)
    except ValueError as error:
        return error


def test_estimates_equal_reference_on_golden_gathers(golden):
    """The reference's estimates of the golden flat-layer gathers (FP32 and FP64),
    recorded by tests/golden/make_surface_golden.py, reproduced exactly by the host
    estimator on the same gathers."""
    ref = _surface_golden()['golden']
    for tag in ('f32', 'f64'):
        obs = golden[f'obs_{tag}'][0]
        assert _digest(obs) == ref[tag]['gathers_sha256']
        _same_estimate(_host_estimate(obs.astype(np.float64)), ref[tag]['estimate'])


def test_estimates_equal_reference_on_its_test_models(observed):
    """The five models of the reference's surface-velocity tests: the oracle's FP32
    gathers are the reference's bytes (SHA-256), and the host estimator returns the
    reference's dicts exactly on them as FP32 and as float64 (and its error message
    for the unresolved refractor)."""
    ref = _surface_golden()['reference_tests']
    for name, obs in zip(NAMES, observed):
        assert _digest(obs) == ref[name]['gathers_sha256'], name
        _same_estimate(_host_estimate(obs), ref[name]['estimate'])
        _same_estimate(_host_estimate(obs.astype(np.float64)), ref[name]['estimate_f64'])
