"""Differentiation workload on the GPU: autograd around the HIP forward, exact
adjoint with checkpoints every 50 steps.

B = 16 seeded mixed models (true) and their Gaussian-smoothed (sigma 8 cells)
starting models, 5 shots, nt = 1000.  Checks: parity with the reference
gradient (golden fixture), the FP64 dot-product test <y, F x> = <F* y, x> with
Gaussian x, y (1e-12), a central finite-difference check on 20 random pixels
in FP64, and checkpoint-interval invariance.
"""
import numpy as np
import pytest
import torch

from shotsweep_amd import autograd as AG
from shotsweep_amd import gpu, synth

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda')


def case(name):
    print(f'\n[case] {name}', flush=True)


@pytest.fixture(scope='module')
def models():
    true = synth.mixed(16, seed=7)
    return true, synth.smooth(true, 8.0)


def test_autograd_gradient_matches_reference(golden):
    case('autograd misfit gradient vs reference gradient (FP32, checkpoints every 50)')
    sim = gpu.Sim(device='cuda', checkpoint_interval=50)
    v = torch.tensor(golden['backgrounds'][:1], device=DEV, dtype=torch.float32, requires_grad=True)
    obs = torch.tensor(golden['obs_f32'], device=DEV)
    loss = AG.misfit(v, obs, sim)
    loss.backward()
    want = golden['grad_f32'].astype(np.float64).sum(1)
    got = v.grad.cpu().numpy().astype(np.float64)
    e = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f'  rel L2 vs reference gradient: {e:.3e}')
    assert e < 1e-5


def test_dot_product_fp64_config2(models):
    case('dot product <y,Fx> = <F*y,x>, FP64, 16 models x 5 shots, nt=1000')
    _, start = models
    sim = gpu.Sim(device='cuda', dtype=torch.float64, checkpoint_interval=50)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(start.shape)
    y = torch.tensor(rng.standard_normal((16, 5, 1000, 70)), device=DEV)
    Fx = sim.born(start, x)
    v = torch.tensor(start, device=DEV, requires_grad=True)
    out = AG.modeling(v, sim)
    out.backward(y)
    lhs = float((y * Fx).sum())
    rhs = float((v.grad * torch.tensor(x, device=DEV)).sum())
    err = abs(lhs - rhs) / abs(lhs)
    print(f'  <y,Fx>={lhs:.15e} <F*y,x>={rhs:.15e} rel {err:.3e}')
    assert err < 1e-12


def test_finite_differences_20_pixels_fp64(models):
    case('central finite differences on 20 random pixels, FP64')
    true, start = models
    sim = gpu.Sim(device='cuda', dtype=torch.float64, checkpoint_interval=50)
    obs = torch.tensor(sim.gathers(true), device=DEV)
    _, grad = AG.misfit_and_gradient(start, obs, sim)
    grad = grad.cpu().numpy()
    rng = np.random.default_rng(3)
    eps = 0.05
    worst = 0.0
    for _ in range(20):
        b, z, x = rng.integers(0, 16), rng.integers(0, 70), rng.integers(0, 70)

        def loss(delta):
            m = start[b:b + 1].copy()
            m[0, z, x] += delta
            r = sim.gathers(m) - obs[b:b + 1].cpu().numpy()
            return 0.5 * float(np.sum(r * r))
        num = (loss(eps) - loss(-eps)) / (2 * eps)
        err = abs(num - grad[b, z, x]) / max(abs(num), 1e-8 * np.abs(grad).max())
        worst = max(worst, err)
    print(f'  worst relative FD error over 20 pixels: {worst:.3e}')
    assert worst < 1e-5


def test_checkpoint_interval_invariance_fp32(models):
    case('checkpoints every 50 vs 73 vs full history, FP32')
    true, start = models
    obs = torch.tensor(gpu.Sim(device='cuda').gathers(true[:4]), device=DEV)
    grads = []
    for ci, hist in ((50, 'checkpoint'), (73, 'checkpoint'), (50, 'full')):
        sim = gpu.Sim(device='cuda', checkpoint_interval=ci, history=hist)
        grads.append(AG.misfit_and_gradient(start[:4], obs, sim)[1].cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])
    np.testing.assert_array_equal(grads[0], grads[2])


def test_misfit_decreases_along_negative_gradient(models):
    case('misfit decreases along -gradient')
    true, start = models
    sim = gpu.Sim(device='cuda', checkpoint_interval=50)
    obs = torch.tensor(sim.gathers(true[:2]), device=DEV)
    f0, g = AG.misfit_and_gradient(start[:2], obs, sim)
    step = 20.0 / float(g.abs().max())
    f1, _ = AG.misfit_and_gradient(start[:2] - step * g.cpu().numpy(), obs, sim)
    print(f'  misfit {f0:.4e} -> {f1:.4e}')
    assert f1 < f0
