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


@pytest.mark.parametrize('N,Cin,Cout,H,W', [
    (40, 32, 32, 70, 70),   # 1400 two-row bands: several per workgroup
    (9, 2, 32, 37, 53),     # first layer, odd sizes (a pad pixel in every band)
    (5, 32, 1, 70, 70),     # last layer
    (3, 8, 40, 12, 10),     # more than 32 output channels: the VALU kernel
])
def test_weight_gradient_matches_float64_and_is_deterministic(N, Cin, Cout, H, W):
    """ss_conv3x3_wgrad_f32 against the FP64 weight and bias gradients of a 3x3
    'same' convolution (torch.nn.grad.conv2d_weight on the CPU); a second call
    gives the same bits."""
    case(f'conv3x3 weight gradient N={N} Cin={Cin} Cout={Cout} {H}x{W}')
    from shotsweep_amd import _lib
    g = torch.Generator().manual_seed(N * 1000 + Cin)
    x = torch.randn(N, Cin, H, W, generator=g)
    dz = torch.randn(N, Cout, H, W, generator=g)
    want = torch.nn.grad.conv2d_weight(x.double(), (Cout, Cin, 3, 3), dz.double(), padding=1)
    want_b = dz.double().sum((0, 2, 3))
    xd, dzd = x.cuda(), dz.cuda()
    outs = []
    for _ in range(2):
        dw = torch.full((Cout, Cin, 3, 3), float('nan'), device='cuda')
        db = torch.full((Cout,), float('nan'), device='cuda')
        _lib.check(_lib.load().ss_conv3x3_wgrad_f32(N, Cin, Cout, H, W, _lib.ptr(xd), _lib.ptr(dzd), _lib.ptr(dw),
                                                   _lib.ptr(db), torch.cuda.current_stream().cuda_stream),
                   'conv3x3_wgrad')
        torch.cuda.synchronize()
        outs.append((dw.cpu(), db.cpu()))
    ew = float((outs[0][0].double() - want).norm() / want.norm())
    eb = float((outs[0][1].double() - want_b).norm() / want_b.norm())
    print(f'  weights rel {ew:.2e}, bias rel {eb:.2e}')
    assert ew < 1e-5 and eb < 1e-5  # FP32 sums of up to 196k products
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
