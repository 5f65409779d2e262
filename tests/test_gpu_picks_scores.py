"""First-break picks and OpenFWI scores on the GPU against the host/oracle versions.

Picks (onset index, peak) are integer/selection work and must equal the
host NumPy picks exactly; the surface-velocity estimates built from them must
then be identical.  Scores are float64 reductions: within 1e-12 relative of
the oracle (torch conv2d, float64).
"""
import numpy as np
import pytest
import torch

from oracle import scores as oscores
from shotsweep_amd import background as BG, gpu, scores, synth

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda')


def case(name):
    print(f'\n[case] {name}', flush=True)


def test_device_picks_equal_host_picks(golden):
    case('first breaks: golden gathers and 12 mixed models, FP32 and FP64')
    sim = gpu.Sim(device=DEV)
    models = synth.mixed(12, seed=21)
    obs = torch.cat([torch.as_tensor(golden['obs_f32'], device=DEV).float(), sim.forward(models)[0]]).contiguous()
    for dtype in (torch.float32, torch.float64):
        o = obs.to(dtype)
        dev = BG.estimate_surface_velocity_device(o)
        host = o.cpu().numpy().astype(np.float64)
        for k in range(len(host)):
            try:
                want = BG.estimate_surface_velocity(host[k])
            except ValueError as e:
                want = e
            got = dev[k]
            if isinstance(want, ValueError):
                assert isinstance(got, ValueError) and str(got) == str(want), k
            else:
                assert got == want, (k, got, want)
        # the raw picks themselves
        on = torch.empty(len(o), 5, 70, dtype=torch.int32, device=DEV)
        pk = torch.empty(len(o), 5, 70, dtype=dtype, device=DEV)
        from shotsweep_amd import _lib
        _lib.check(_lib.fn('first_breaks', dtype)(len(o) * 5, 1000, 70, _lib.ptr(o), 0.05, _lib.ptr(on),
                                                  _lib.ptr(pk), _lib.stream_handle(DEV)), 'first_breaks')
        for k in range(len(host)):
            onset, peaks = BG.first_breaks(host[k])
            np.testing.assert_array_equal(on[k].cpu().numpy(), onset)
            np.testing.assert_array_equal(pk[k].cpu().numpy().astype(np.float64), peaks)


@pytest.mark.parametrize('shape', [(5, 70, 70), (3, 37, 53)])
def test_scores_match_oracle(shape):
    case(f'velocity scores {shape}')
    true = synth.mixed(shape[0], seed=7, nz=shape[1], nx=shape[2])
    pred = np.clip(synth.smooth(true, 3.0) + np.random.default_rng(1).normal(0, 40, true.shape), 1500, 4500)
    for dtype in (torch.float64, torch.float32):
        p = torch.tensor(pred, dtype=dtype, device=DEV)
        t = torch.tensor(true, dtype=dtype, device=DEV)
        got = scores.velocity_scores(p, t).cpu().numpy()
        for k in range(shape[0]):
            want = oscores.metrics(p[k].double().cpu().numpy(), t[k].double().cpu().numpy())
            np.testing.assert_allclose(got[k], [want['mae'], want['rmse'], want['ssim']], rtol=1e-12, atol=0)
    same = scores.velocity_scores(t, t).cpu().numpy()
    assert np.all(same[:, :2] == 0) and np.allclose(same[:, 2], 1.0, rtol=0, atol=1e-14)
    d = scores.metrics(pred[0], true[0])
    w = oscores.metrics(pred[0], true[0])
    assert set(d) == {'mae', 'rmse', 'ssim'} and all(abs(d[k] - w[k]) <= 1e-12 * abs(w[k]) for k in d)


def test_device_estimates_equal_reference_values():
    """GPU gathers of the reference's five surface-velocity test models are the
    reference's bytes, and the GPU picks plus host fits on them as float64 (as
    the harness reads gathers) return the reference's recorded float64 estimates
    exactly (tests/golden/make_surface_golden.py)."""
    from test_background import NAMES, _digest, _same_estimate, _surface_golden
    print('\n[case] surface-velocity estimates vs reference values', flush=True)
    ref = _surface_golden()['reference_tests']
    models = [np.full((70, 70), v, dtype=float) for v in (1500, 2500, 4500)]
    for depth in (5, 2):
        m = np.full((70, 70), 1500.0)
        m[depth:] = 4500.0
        models.append(m)
    obs = gpu.Sim(device='cuda').forward(np.array(models))[0].contiguous()
    host = obs.cpu().numpy()
    for name, o in zip(NAMES, host):
        assert _digest(o) == ref[name]['gathers_sha256'], name
    for name, est in zip(NAMES, BG.estimate_surface_velocity_device(obs.double())):
        _same_estimate(est, ref[name]['estimate_f64'])
