"""The reference's own physics tests (tests/test_physics.py:25-88 of the reference),
restated against shotsweep_amd.gpu with the reference's call patterns: a Sim on
device='cpu' in FP64 (the kernels run on the GPU, results come back to the CPU),
gpu.lap / gpu.lap_adjoint on CPU tensors, finite differences along the
interior, the left, bottom and top boundary strips and the velocity-dependent
source pixel, the muted finite difference on row 69, and checkpoint-interval
invariance at interval 73.  The checks are the reference's; the tolerances are
its own (FD relative error < 1e-4; rtol 1e-10 for the checkpoint invariance)."""
import numpy as np
import pytest
import torch

from oracle import fd
from shotsweep_amd import gpu

pytestmark = pytest.mark.gpu


def case(name):
    print(f'\n[case] {name}', flush=True)


@pytest.fixture(scope='module')
def adjoint_case():
    sim = gpu.Sim(device='cpu', dtype=torch.float64)
    velocity = np.full((1, 70, 70), 2000.0)
    truth = velocity.copy()
    truth[:, 8:] = 2400.0
    return sim, velocity, sim.gathers(truth)


def test_cpu_device_call_surface(adjoint_case):
    """Sim(device='cpu') returns what the reference's CPU Sim returns: NumPy gathers,
    forward() tensors on the CPU with c2 shaped (B, 1, H, W), checkpoints on the CPU,
    mute weights on the CPU (test_physics.py:68 calls .cpu() on them)."""
    case("Sim(device='cpu', dtype=float64) call surface")
    sim, velocity, observed = adjoint_case
    assert sim.device == torch.device('cpu') and sim.kdev.type == 'cuda'
    assert isinstance(observed, np.ndarray) and observed.dtype == np.float64 and observed.shape == (1, 5, 1000, 70)
    out, ck, c2 = sim.forward(velocity, store=True)
    assert out.device.type == 'cpu' and c2.device.type == 'cpu' and tuple(c2.shape) == (1, 1, 150, 150)
    assert all(p.device.type == 'cpu' for pair in ck.values() for p in pair) and sorted(ck)[:2] == [0, 40]
    np.testing.assert_array_equal(out.numpy(), sim.gathers(velocity))
    w = sim.mute_weights(np.array([2000.0]))
    assert w.device.type == 'cpu' and w.dtype == torch.float64 and tuple(w.shape) == (1, 5, 1000, 70)
    # the taper's cosine: correctly rounded here, torch's CPU cosine in the oracle (a few ulp apart)
    np.testing.assert_allclose(w.cpu().numpy(), fd.OracleSim(dtype=torch.float64).mute_weights(
        np.array([2000.0])).numpy(), rtol=0, atol=4 * np.finfo(np.float64).eps)


def test_laplacian_transpose_including_boundary_cpu_tensors():
    """test_physics.py:35-39 on CPU tensors: <lap x, y> = <x, lap^T y> to 10 places, and
    both operators equal the oracle's to the bit; the results come back on the CPU."""
    case('lap / lap_adjoint on CPU tensors')
    generator = torch.Generator().manual_seed(8)
    x = torch.randn(2, 9, 11, generator=generator, dtype=torch.float64)
    y = torch.randn(2, 9, 11, generator=generator, dtype=torch.float64)
    lx, lty = gpu.lap(x), gpu.lap_adjoint(y)
    assert lx.device.type == 'cpu' and lty.device.type == 'cpu' and lx.dtype == torch.float64
    assert abs(float((lx * y).sum()) - float((x * lty).sum())) < 0.5e-10
    np.testing.assert_array_equal(lx.numpy(), fd.lap(x).numpy())
    np.testing.assert_array_equal(lty.numpy(), fd.lap_adjoint(y).numpy())
    g = torch.randn(2, 3, 19, 23, generator=generator, dtype=torch.float64)
    f = gpu.fold_edge_padding(g, 5)
    assert f.device.type == 'cpu'
    np.testing.assert_allclose(f.numpy(), fd.fold_edge_padding(g, 5, 9, 13).numpy(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('name,location', [
    ('interior', (slice(10, 30), 35)),
    ('left_boundary', (slice(10, 30), 0)),
    ('bottom_boundary', (69, slice(25, 45))),
    ('top_boundary', (0, slice(25, 45))),
    ('velocity_dependent_source', (1, 34)),
])
def test_source_and_padding_derivatives_match_finite_differences(adjoint_case, name, location):
    """test_physics.py:41-61: the adjoint gradient summed over shots against a central
    difference (epsilon 0.1 m/s) of the misfit along directions covering the
    interior, the padding folds of the left, bottom and top strips and the
    velocity-dependent source amplitude at (1, 34)."""
    case(f'FD direction {name} FP64 70x70')
    sim, velocity, observed = adjoint_case
    gradient = sim.gradient(velocity, observed)[0].sum(1)
    direction = np.zeros_like(velocity)
    direction[(0,) + location] = 1
    epsilon = 0.1

    def loss(v):
        residual = sim.gathers(v) - observed
        return 0.5 * np.sum(residual * residual)
    numerical = (loss(velocity + epsilon * direction) - loss(velocity - epsilon * direction)) / (2 * epsilon)
    adjoint = np.sum(gradient * direction)
    print(f'  numerical {numerical:.6e} adjoint {adjoint:.6e}')
    assert abs(numerical - adjoint) / max(abs(numerical), 1.0) < 1e-4


def test_muted_gradient_and_checkpoint_recomputation(adjoint_case):
    """test_physics.py:63-79: the muted gradient against a finite difference along row
    69, columns 25-44 (mute weights held fixed), then a Sim with checkpoints every 73
    steps gives the same gradient (rtol 1e-10) and illumination (rtol 1e-12)."""
    case('muted FD on row 69 and checkpoint interval 73, FP64')
    sim, velocity, observed = adjoint_case
    mute = np.array([2000.0])
    gradient, illumination = sim.gradient(velocity, observed, v_mute=mute)
    direction = np.zeros_like(velocity)
    direction[:, 69, 25:45] = 1
    weights = sim.mute_weights(mute).cpu().numpy()

    def loss(v):
        residual = sim.gathers(v) - observed
        return 0.5 * np.sum(weights * residual ** 2)
    epsilon = 0.1
    numerical = (loss(velocity + epsilon * direction) - loss(velocity - epsilon * direction)) / (2 * epsilon)
    adjoint = (gradient.sum(1) * direction).sum()
    assert abs(numerical - adjoint) / abs(numerical) < 1e-4
    other = gpu.Sim(device='cpu', dtype=torch.float64, checkpoint_interval=73, history='checkpoint')
    other_gradient, other_illumination = other.gradient(velocity, observed, v_mute=mute)
    np.testing.assert_allclose(other_gradient, gradient, rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(other_illumination, illumination, rtol=1e-12, atol=1e-6)
