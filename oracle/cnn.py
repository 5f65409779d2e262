"""CPU reference of the learned-preconditioner CNN (test infrastructure only).

The same network as shotsweep_amd.cnn (3x3 same convolutions 2-32-32-32-1,
ReLU after the first three), written with torch.nn.functional on the CPU in
float64, with autograd for the parameter gradients and a plain Adam step.
No reference counterpart exists in openfwi (refinement is not implemented
there); the checks pin the hand-written kernels to this textbook definition.
"""
import torch
import torch.nn.functional as F


def unflatten(flat, layers):
    out, off = [], 0
    for ci, co, relu in layers:
        nw = co * ci * 9
        w = flat[off:off + nw].view(co, ci, 3, 3)
        b = flat[off + nw:off + nw + co]
        out.append((w, b, relu))
        off += nw + co
    return out


def forward(flat, x, layers):
    h = x
    for w, b, relu in unflatten(flat, layers):
        h = F.conv2d(h, w, b, padding=1)
        if relu:
            h = torch.relu(h)
    return h


def loss_and_grad(flat, x, target, layers):
    """0.5*mean((net(x) - target)^2) and d/dflat, float64 on the CPU."""
    p = flat.detach().double().clone().requires_grad_(True)
    y = forward(p, x.double(), layers)
    loss = 0.5 * ((y - target.double()) ** 2).mean()
    loss.backward()
    return float(loss), p.grad.detach()


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / ((v / (1 - b2 ** step)).sqrt() + eps)
    return p, m, v
