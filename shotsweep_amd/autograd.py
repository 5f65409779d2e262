"""Differentiable forward modelling: torch.autograd around the HIP simulator.

``modeling(v, sim)`` returns the gathers (B, ns, nt, nx) of velocity models
v (B, nz, nx); its backward is the hand-written exact discrete adjoint
(``Sim._adjoint_tape``, one native ss_rtm call), seeded with the incoming
gradient, recomputing the forward from checkpoints every
``sim.checkpoint_interval`` steps (or reading the stored Laplacian history
when the Sim has ``history='full'``).  For
``0.5 * ((modeling(v) - obs)**2).sum()`` this is the reference's misfit
gradient (openfwi/physics/gpu.py:152-194, summed over shots).

No reference counterpart for the autograd wrapper itself: the reference calls
its adjoint directly (gpu.py:190-194).
"""
import numpy as np
import torch


class _Modeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, sim):
        prep = sim._prepare(v.detach())   # coefficients formed on the device, no host round trip
        keep = v.requires_grad
        # the tape: checkpoints every checkpoint_interval steps with recompute,
        # unless the Sim was built with history='full'
        seg = sim.geom.nt if (not keep or sim.history == 'full') else sim.checkpoint_interval
        gathers, tape = sim._forward_tape(prep, segment=seg, fresh=True)
        if not keep:
            tape = None
        ctx.sim, ctx.prep, ctx.tape = sim, prep, tape
        ctx.v_meta = (v.dtype, v.device)
        return gathers

    @staticmethod
    def backward(ctx, grad_out):
        sim, tape = ctx.sim, ctx.tape
        if tape is None:
            raise RuntimeError('modeling() ran without requires_grad; nothing was taped')
        res = grad_out.to(sim.device, sim.dtype).contiguous()
        grad, _ = sim._adjoint_tape(ctx.prep, tape, res)
        ctx.tape = None
        dtype, device = ctx.v_meta
        return grad.sum(1).to(device, dtype), None


def modeling(v, sim):
    """Gathers of v through the HIP simulator, differentiable in v."""
    return _Modeling.apply(v, sim)


def misfit(v, obs, sim):
    """0.5 * sum((gathers(v) - obs)^2) as a differentiable scalar."""
    r = modeling(v, sim) - obs
    return 0.5 * (r * r).sum()


def misfit_and_gradient(v, obs, sim):
    """(misfit, dmisfit/dv) for numpy or tensor inputs; gradient as a tensor like v."""
    v = torch.as_tensor(np.asarray(v) if not torch.is_tensor(v) else v, dtype=sim.dtype, device=sim.device)
    v = v.detach().requires_grad_(True)
    obs = torch.as_tensor(obs, dtype=sim.dtype, device=sim.device)
    loss = misfit(v, obs, sim)
    loss.backward()
    return float(loss.detach()), v.grad.detach()
