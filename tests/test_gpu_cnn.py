"""Hand-written CNN kernels of the learned preconditioner vs the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import cnn as ocnn
from shotsweep_amd.cnn import LAYERS, PrecondCNN, n_params

pytestmark = pytest.mark.gpu


def case(name):
    print(f'\n[case] {name}', flush=True)


def test_parameter_count():
    assert n_params() == 19393


def test_forward_backward_match_oracle():
    case('CNN forward / parameter gradient vs torch CPU float64')
    net = PrecondCNN('cuda', seed=1)
    g = torch.Generator().manual_seed(2)
    x = torch.randn(3, 2, 70, 70, generator=g)
    t = torch.randn(3, 1, 70, 70, generator=g) * 0.1
    y, acts = net.forward(x.cuda(), keep=True)
    flat = net.params.cpu()
    want = ocnn.forward(flat.double(), x.double(), LAYERS)
    ef = float((y.cpu().double() - want).norm() / want.norm())
    loss, dy = net.loss_grad(y, t.cuda())
    grads = net.backward(acts, dy).cpu().double()
    oloss, ograd = ocnn.loss_and_grad(flat, x, t, LAYERS)
    eg = float((grads - ograd).norm() / ograd.norm())
    print(f'  forward rel {ef:.2e}; loss {loss:.6e} vs {oloss:.6e}; grad rel {eg:.2e}')
    assert ef < 1e-5 and abs(loss - oloss) / oloss < 1e-5 and eg < 1e-4


def test_ragged_size_and_adam_step():
    case('CNN on 37x53 inputs; Adam step vs oracle')
    net = PrecondCNN('cuda', seed=3)
    g = torch.Generator().manual_seed(4)
    x = torch.randn(2, 2, 37, 53, generator=g)
    t = torch.randn(2, 1, 37, 53, generator=g) * 0.1
    p0 = net.params.cpu().double()
    loss = net.train_step(x.cuda(), t.cuda(), lr=1e-3)
    _, ograd = ocnn.loss_and_grad(p0.float(), x, t, LAYERS)
    p1, _, _ = ocnn.adam(p0, ograd, torch.zeros_like(p0), torch.zeros_like(p0), 1)
    # the first Adam step is ~ lr * sign(g): components whose FP32 and FP64 gradients
    # straddle zero may flip, so count the components that differ by more than 1 % of lr
    off = ((net.params.cpu().double() - p1).abs() > 1e-5).double().mean().item()
    print(f'  loss {loss:.4e}; Adam components off by > 1% of lr: {off:.2e}')
    assert off < 1e-3


def test_training_reduces_loss():
    case('a few hundred steps fit a fixed batch')
    net = PrecondCNN('cuda', seed=5)
    g = torch.Generator().manual_seed(6)
    x = torch.randn(8, 2, 70, 70, generator=g).cuda()
    t = (0.3 * x[:, :1] - 0.2 * x[:, 1:]).contiguous()
    first = net.train_step(x, t, lr=3e-3)
    for _ in range(200):
        last = net.train_step(x, t, lr=3e-3)
    print(f'  loss {first:.4e} -> {last:.4e}')
    assert last < 0.1 * first
