"""Test-time refinement with a learned preconditioner vs the plain adjoint-gradient step.

Workload (no reference counterpart: refinement is "not implemented yet" in the
reference, README.md:25; the plain step with backtracking follows the
archive's exact-adjoint FWI, archive/toy/slopes/fwi.py:138-164):

* training pairs: seeded mixed-structure models (the batched-RTM generator),
  current models = Gaussian smoothing of the truth with sigma drawn from
  [2, 10] cells; input = (shot-summed misfit gradient, shot-summed
  illumination) of the current model against the true gathers, each
  normalised by its maximum; target = (true - current) / 1000 (km/s);
* network: ``PrecondCNN`` trained with Adam on the hand-written kernels, then
  dataset aggregation: pairs along the trained network's own refinement
  trajectories are added and a fresh network is trained on all of them (twice
  by default), so it learns the models its own updates produce;
* refinement: start = smooth(true, 8); each iteration computes the gradient
  (HIP forward + exact adjoint) and
    plain:   d = -g/max|g|, backtracking on the data misfit from a 100 m/s
             maximum update, halving up to 6 times;
    learned: d = 1000 * CNN(features), steps 1, 1/2, 1/4 of it; a model whose
             misfit does not decrease falls back to the plain step;
  a model keeps its velocity when no trial decreases its misfit.

    python -m shotsweep_amd.refine --pairs 20000 --epochs 8 --eval 256 --iters 10 --out result.json
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _lib, gpu, scores, synth
from .cnn import PrecondCNN

VMIN, VMAX = 1500.0, 4500.0
PLAIN_STEP = 100.0          # m/s, largest plain update before backtracking
PLAIN_HALVINGS = 6
LEARNED_STEPS = (1.0, 0.5, 0.25)


class Problem:
    """A batch of models with fixed observed gathers on the device."""

    def __init__(self, sim, true_models):
        self.sim = sim
        self.obs = sim.forward(true_models)[0].contiguous().clone()
        self.B = len(true_models)
        g = sim.geom
        self.nzx = g.nz * g.nx
        self.ntr = g.ns * g.nt * g.nx

    def misfit(self, v):
        """Per-model data misfit 0.5*sum((F(v) - obs)^2), v a (B, nz, nx) device tensor."""
        sim = self.sim
        gathers = sim.forward(v.detach())[0]
        out = torch.empty(self.B, dtype=torch.float32, device=sim.device)
        _lib.check(_lib.load().ss_sumsq_diff_f32(self.B, self.ntr, _lib.ptr(gathers), _lib.ptr(self.obs),
                                                 _lib.ptr(out), sim._stream()), 'sumsq_diff')
        return 0.5 * out

    def features(self, v):
        """CNN input (B, 2, nz, nx) and per-model misfit at v (one forward + adjoint)."""
        sim = self.sim
        g = sim.geom
        prep = sim._prepare(v.detach())
        grad, illum, gathers = sim._gradient_device(prep, self.obs, 1, None)
        x = torch.empty(self.B, 2, g.nz, g.nx, dtype=torch.float32, device=sim.device)
        gscale = torch.empty(self.B, dtype=torch.float32, device=sim.device)
        lib = _lib.load()
        _lib.check(lib.ss_precond_features_f32(self.B, g.ns, self.nzx, _lib.ptr(grad), _lib.ptr(illum), _lib.ptr(x),
                                               _lib.ptr(gscale), sim._stream()), 'precond_features')
        f = torch.empty(self.B, dtype=torch.float32, device=sim.device)
        _lib.check(lib.ss_sumsq_diff_f32(self.B, self.ntr, _lib.ptr(gathers), _lib.ptr(self.obs), _lib.ptr(f),
                                         sim._stream()), 'sumsq_diff')
        return x, 0.5 * f

    def update(self, v, d, alpha):
        """clamp(v + alpha[b] * d, VMIN, VMAX) for every model b."""
        out = torch.empty_like(v)
        _lib.check(_lib.load().ss_velocity_update_f32(self.B, self.nzx, _lib.ptr(v.contiguous()),
                                                      _lib.ptr(d.contiguous()), _lib.ptr(alpha.contiguous()), VMIN,
                                                      VMAX, _lib.ptr(out), self.sim._stream()), 'velocity_update')
        return out


def _search(prob, v, d, f0, steps, active):
    """Backtracking over ``steps`` for the models in ``active``: the first step that lowers the misfit wins."""
    B = prob.B
    best_v, best_f = v.clone(), f0.clone()
    done = ~active
    for s in steps:
        if bool(done.all()):
            break
        alpha = torch.where(done, torch.zeros_like(f0), torch.full_like(f0, s))
        trial = prob.update(v, d, alpha)
        ft = prob.misfit(trial)
        win = (~done) & (ft < f0)
        best_v[win] = trial[win]
        best_f[win] = ft[win]
        done = done | win
    return best_v, best_f, done & active


def plain_step(prob, v, x, f0, active=None):
    active = torch.ones(prob.B, dtype=torch.bool, device=v.device) if active is None else active
    d = -x[:, 0]
    steps = [PLAIN_STEP * 0.5 ** k for k in range(PLAIN_HALVINGS + 1)]
    return _search(prob, v, d, f0, steps, active)


def learned_step(prob, net, v, x, f0):
    d = 1000.0 * net(x)[:, 0]
    v1, f1, ok = _search(prob, v, d, f0, LEARNED_STEPS, torch.ones(prob.B, dtype=torch.bool, device=v.device))
    fallback = ~ok
    if bool(fallback.any()):
        v2, f2, _ = plain_step(prob, v, x, f0, fallback)
        v1[fallback] = v2[fallback]
        f1[fallback] = f2[fallback]
    return v1, f1, int(fallback.sum())


def refine(prob, start, iters, net=None, log=None):
    """Misfit history (B, iters+1), model history RMS error and fallback counts."""
    v = torch.as_tensor(start, dtype=torch.float32, device=prob.sim.device).contiguous()
    hist, falls = [], []
    for k in range(iters):
        x, f = prob.features(v)
        hist.append(f.cpu().numpy())
        if net is None:
            v, _, _ = plain_step(prob, v, x, f)
            falls.append(0)
        else:
            v, _, nf = learned_step(prob, net, v, x, f)
            falls.append(nf)
        if log:
            log(f'  iter {k + 1}: mean misfit ratio {float(np.mean(hist[-1] / hist[0])):.4f}')
    hist.append(prob.misfit(v).cpu().numpy())
    return np.stack(hist, 1), v, falls


def make_pairs(sim, n_pairs, seed=0, batch=128, log=None):
    """Training inputs (n, 2, nz, nx) and targets (n, 1, nz, nx) on the device."""
    g = sim.geom
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    made = 0
    while made < n_pairs:
        b = min(batch, n_pairs - made)
        true = synth.mixed(b, seed=int(rng.integers(1 << 30)), nz=g.nz, nx=g.nx)
        sig = rng.uniform(2.0, 10.0, b)
        cur = np.stack([synth.smooth(true[k:k + 1], sig[k])[0] for k in range(b)])
        prob = Problem(sim, true)
        x, _ = prob.features(torch.tensor(cur, dtype=torch.float32, device=sim.device))
        xs.append(x)
        ts.append(torch.tensor((true - cur)[:, None] / 1000.0, dtype=torch.float32, device=sim.device))
        made += b
        if log:
            log(f'  pairs {made}/{n_pairs}')
    return torch.cat(xs), torch.cat(ts)


def trajectory_pairs(sim, net, n_models, steps, seed=0, batch=128, log=None):
    """Training pairs along the learned refinement's own trajectories: for seeded
    true models and smooth starts (sigma in [2, 10] cells), the features and target
    (true - current) / 1000 at the start and after each of ``steps`` learned steps
    (with their fallback).  Adding these to the training set and training again
    (dataset aggregation) teaches the network the models its own updates produce,
    not only smooth starts."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    made = 0
    while made < n_models:
        b = min(batch, n_models - made)
        true = synth.mixed(b, seed=int(rng.integers(1 << 30)), nz=sim.geom.nz, nx=sim.geom.nx)
        sig = rng.uniform(2.0, 10.0, b)
        cur = np.stack([synth.smooth(true[k:k + 1], sig[k])[0] for k in range(b)])
        prob = Problem(sim, true)
        tt = torch.tensor(true, dtype=torch.float32, device=sim.device)
        v = torch.tensor(cur, dtype=torch.float32, device=sim.device)
        for k in range(steps + 1):
            x, f = prob.features(v)
            xs.append(x)
            ts.append(((tt - v) / 1000.0)[:, None].contiguous())
            if k < steps:
                v, _, _ = learned_step(prob, net, v, x, f)
        made += b
        if log:
            log(f'  trajectory models {made}/{n_models}')
    return torch.cat(xs), torch.cat(ts)


def train(net, x, t, epochs, batch=64, lr=1e-3, seed=0, log=None):
    n = len(x)
    g = torch.Generator().manual_seed(seed)
    losses = []
    steps = epochs * ((n + batch - 1) // batch)
    k = 0
    for ep in range(epochs):
        order = torch.randperm(n, generator=g).to(x.device)
        tot = 0.0
        for i in range(0, n, batch):
            idx = order[i:i + batch]
            lr_k = lr * 0.5 * (1 + np.cos(np.pi * k / steps))        # cosine decay
            tot += net.train_step(x[idx].contiguous(), t[idx].contiguous(), lr=lr_k) * len(idx)
            k += 1
        losses.append(tot / n)
        if log:
            log(f'  epoch {ep + 1}: loss {losses[-1]:.4e}')
    return losses


def iterations_to_reach(hist, target):
    """First iteration whose mean misfit ratio is <= target (None if never)."""
    ratio = (hist / hist[:, :1]).mean(0)
    hit = np.nonzero(ratio <= target)[0]
    return int(hit[0]) if len(hit) else None


def run(pairs=20000, epochs=8, n_eval=256, iters=10, seed=0, dagger_rounds=2, traj_steps=3, log=print):
    """Generate pairs, train, then ``dagger_rounds`` times add ``pairs`` more pairs
    from the current network's own refinement trajectories (pairs / (traj_steps + 1)
    models, traj_steps learned steps each) and train a fresh network on everything;
    evaluate the last network against the plain step on held-out models."""
    sim = gpu.Sim(device='cuda')
    t0 = time.time()
    log(f'[refine] generating {pairs} training pairs')
    x, t = make_pairs(sim, pairs, seed=seed, log=log)
    t_pairs = time.time() - t0
    net = PrecondCNN('cuda', seed=seed)
    log(f'[refine] training {epochs} epochs')
    t1 = time.time()
    losses = train(net, x, t, epochs, log=log)
    t_train = time.time() - t1
    for r in range(dagger_rounds):
        t2 = time.time()
        log(f'[refine] aggregation round {r + 1}: trajectories of the current network')
        xa, ta = trajectory_pairs(sim, net, max(1, pairs // (traj_steps + 1)), traj_steps, seed=seed + 1 + r,
                                  log=log)
        x, t = torch.cat([x, xa]), torch.cat([t, ta])
        t_pairs += time.time() - t2
        t3 = time.time()
        net = PrecondCNN('cuda', seed=seed)
        losses = train(net, x, t, epochs, log=log)
        t_train += time.time() - t3
    n_train = len(x)
    del x, t
    true = synth.mixed(n_eval, seed=10_000 + seed)
    start = synth.smooth(true, 8.0)
    prob = Problem(sim, true)
    log('[refine] plain')
    torch.cuda.synchronize()
    t2 = time.time()
    hp, vp, _ = refine(prob, start, iters, None, log)
    torch.cuda.synchronize()
    t_plain = time.time() - t2
    log('[refine] learned')
    t3 = time.time()
    hl, vl, falls = refine(prob, start, iters, net, log)
    torch.cuda.synchronize()
    t_learned = time.time() - t3
    plain_ratio = (hp / hp[:, :1]).mean(0)
    learned_ratio = (hl / hl[:, :1]).mean(0)
    tt = torch.tensor(true, dtype=torch.float32, device=sim.device)
    ts = torch.tensor(start, dtype=torch.float32, device=sim.device)

    def openfwi(v):   # mean OpenFWI MAE / RMSE / SSIM over the held-out models
        m = scores.velocity_scores(v.contiguous(), tt).mean(0).tolist()
        return dict(zip(('mae', 'rmse', 'ssim'), m))

    result = {
        'pairs': pairs, 'training_pairs': n_train, 'dagger_rounds': dagger_rounds, 'traj_steps': traj_steps,
        'epochs': epochs, 'eval_models': n_eval, 'iters': iters, 'seed': seed,
        'train_loss': losses, 'time_pairs_s': t_pairs, 'time_train_s': t_train,
        'time_refine_plain_s': t_plain, 'time_refine_learned_s': t_learned,
        'refine_iterations_per_s': n_eval * iters / t_plain,
        'plain_mean_misfit_ratio': plain_ratio.tolist(),
        'learned_mean_misfit_ratio': learned_ratio.tolist(),
        'learned_fallbacks_per_iter': falls,
        'learned_iters_to_reach_plain_final': iterations_to_reach(hl, plain_ratio[-1]),
        'model_rms_error_start': float(((torch.tensor(start, device=sim.device) - tt) ** 2).mean().sqrt()),
        'model_rms_error_plain': float(((vp - tt) ** 2).mean().sqrt()),
        'model_rms_error_learned': float(((vl - tt) ** 2).mean().sqrt()),
        'openfwi_scores_start': openfwi(ts), 'openfwi_scores_plain': openfwi(vp),
        'openfwi_scores_learned': openfwi(vl),
    }
    return result, net


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--pairs', type=int, default=20000)
    ap.add_argument('--epochs', type=int, default=8)
    ap.add_argument('--eval', type=int, default=256)
    ap.add_argument('--iters', type=int, default=10)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--dagger-rounds', type=int, default=2)
    ap.add_argument('--traj-steps', type=int, default=3)
    ap.add_argument('--out', default=None)
    args = ap.parse_args()
    result, net = run(args.pairs, args.epochs, args.eval, args.iters, args.seed, args.dagger_rounds, args.traj_steps,
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(result)
    print(text, flush=True)
    if args.out:
        with open(args.out, 'w') as f:
            f.write(text + '\n')
        np.savez(args.out.replace('.json', '') + '_net.npz', **net.state_dict())


if __name__ == '__main__':
    main()
