"""Mixed-size RTM sweep: models on different grids, step counts and shot counts together.

The reference images one 70 x 70 geometry per ``Sim`` (openfwi/physics/gpu.py:85,
95, 103, 126); it has no counterpart for mixing sizes.  Here every grid size is
a *group* (an ``Acquisition``: nz, nx, nt, ns, source columns), and one sweep
runs all groups' propagations through the tiled kernels of libshotsweep.so:

  1. forward with the whole Laplacian history, every group in one launch per
     step range (``ss_sweep_forward_f32``, replayed from a hipGraph);
  2. muted residual per group (gpu.py:142-160, ``ss_residual_f32``);
  3. exact adjoint with imaging condition, illumination and source term, again
     one launch per step range for all groups (``ss_sweep_adjoint_f32``);
  4. gradient fold and image epilogue per group (gpu.py:185-210).

Per cell the arithmetic is the reference's (the same kernels' helpers as
``gpu.Sim``), so each group's images equal ``gpu.Sim(geom=acq).images`` on the
same inputs bit for bit.  Tile shape and temporal depth per grid size come from
the autotuner (``ss_sweep_autotune_f32``) on the first call.  When the
Laplacian histories of all groups exceed the memory budget, the models are
split into chunks, each a proportional slice of every group.
"""
import ctypes
import itertools
import math

import numpy as np
import torch

from . import _lib, gpu
from .constants import Acquisition

FORWARD, ADJOINT = 0, 1


class MixedSweep:
    """Batched forward + exact-adjoint RTM images over up to four grid sizes."""

    def __init__(self, acquisitions, device='cuda', pad=40, R=1e-5, autotune=True, use_graph=True,
                 max_history_bytes=None, tiles=None):
        acquisitions = list(acquisitions)
        if not 1 <= len(acquisitions) <= _lib.SWEEP_MAX_GROUPS:
            raise ValueError(f'between 1 and {_lib.SWEEP_MAX_GROUPS} grid sizes are supported')
        self.sims = [gpu.Sim(geom=a, pad=pad, R=R, device=device) for a in acquisitions]
        self.acqs = acquisitions
        self.device = self.sims[0].device
        self.autotune = autotune
        self.use_graph = use_graph
        self.max_history_bytes = max_history_bytes
        # per group and direction: (tile_h, tile_w, depth); None = tune or model
        self.tiles = [dict(tiles[k]) if tiles is not None else {} for k in range(len(acquisitions))]
        self._buf = [{} for _ in acquisitions]
        self._plans = {}
        self._rounds = {}
        self.events = None   # list -> (name, start, end) HIP events around the two sweep launches

    # ------------------------------------------------------------- buffers
    def _get(self, k, name, shape):
        n = int(np.prod(shape))
        t = self._buf[k].get(name)
        if t is None or t.numel() < n:
            self._buf[k].pop(name, None)
            t = torch.empty(n, dtype=torch.float32, device=self.device)
            self._buf[k][name] = t
        return t[:n].view(*shape)

    def release(self):
        self._plans.clear()
        for b in self._buf:
            b.clear()
        for s in self.sims:
            s.release()

    def history_bytes(self, rows):
        """Laplacian-history bytes of a sweep over ``rows[k]`` velocity rows of group k."""
        tot = 0
        for sim, nb in zip(self.sims, rows):
            g = sim.geom
            tot += nb * g.ns * g.nt * sim.H * sim.W * 4
        return tot

    def _stream(self):
        return _lib.stream_handle(self.device)

    def _timed(self, name):
        if self.events is None:
            return None
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        self.events.append((name, start, end))
        return end

    # ---------------------------------------------------------------- sweep
    def _groups(self, parts, direction):
        arr = (_lib.SSSweepGroup * len(parts))()
        for i, (k, prep, bufs) in enumerate(parts):
            sim = self.sims[k]
            g = arr[i]
            g.geom = sim._geom(prep['nb'])
            th, tw, d = self.tiles[k].get(direction, (0, 0, 0))
            g.tile_h, g.tile_w, g.depth = th, tw, d
            g.c2 = _lib.ptr(prep['c2'])
            g.A = _lib.ptr(sim.A)
            g.B = _lib.ptr(sim.B)
            g.w = _lib.ptr(sim.w)
            g.amp = _lib.ptr(prep['amp'] if direction == FORWARD else prep['sd'])
            g.state = _lib.ptr(bufs['state'])
            g.hist = _lib.ptr(bufs['hist'])
            if direction == FORWARD:
                g.gathers = _lib.ptr(bufs['sim'])
            else:
                g.res = _lib.ptr(bufs['res'])
                g.gacc = _lib.ptr(bufs['gacc'])
                g.illum = _lib.ptr(bufs['illum'])
                g.src_grad = _lib.ptr(bufs['sg'])
        return arr

    def _run(self, parts, direction):
        lib = _lib.load()
        arr = self._groups(parts, direction)
        n = len(parts)
        if self.autotune and any(direction not in self.tiles[k] for k, _, _ in parts):
            _lib.check(lib.ss_sweep_autotune_f32(n, arr, direction, 0, self._stream()), 'sweep autotune')
            for i, (k, _, _) in enumerate(parts):
                self.tiles[k][direction] = (arr[i].tile_h, arr[i].tile_w, arr[i].depth)
            arr = self._groups(parts, direction)
        f = lib.ss_sweep_forward_f32 if direction == FORWARD else lib.ss_sweep_adjoint_f32
        end = self._timed('sweep_forward' if direction == FORWARD else 'sweep_adjoint')
        _lib.check(f(n, arr, int(self.use_graph), self._stream()), 'sweep')
        if end is not None:
            end.record()

    def _stable_prep(self, k, prep):
        """prep with c2, amp and the source derivative copied into this group's cached
        buffers: the sweeps replay hipGraphs keyed by buffer addresses, and _prepare
        makes new tensors on every call."""
        out = dict(prep)
        for key in ('c2', 'amp', 'sd'):
            buf = self._get(k, 'prep_' + key, tuple(prep[key].shape))
            buf.copy_(prep[key])
            out[key] = buf
        return out

    def _buffers(self, k, nb):
        sim = self.sims[k]
        g = sim.geom
        H, W = sim.H, sim.W
        return {
            'state': self._get(k, 'state', (4, nb, g.ns, H, W)),
            'hist': self._get(k, 'hist', (nb, g.ns, g.nt, H, W)),
            'sim': self._get(k, 'sim', (nb, g.ns, g.nt, g.nx)),
            'res': self._get(k, 'res', (nb, g.ns, g.nt, g.nx)),
            'gacc': self._get(k, 'gacc', (nb, g.ns, H, W)),
            'illum': self._get(k, 'illum', (nb, g.ns, g.nz, g.nx)),
            'sg': self._get(k, 'sg', (nb, g.ns)),
        }

    @torch.no_grad()
    def gradients_device(self, batches, t_mute=0.12):
        """Forward, residual, adjoint and fold for one chunk.

        ``batches[k]`` is None or a dict with ``prep`` (``Sim._prepare`` of the
        rows), ``obs`` (device (nb / obs_repeat, ns, nt, nx)), ``obs_repeat`` and
        ``v_mute`` (device (nb,) or None).  Returns per group (grad, illum, sim
        gathers) device tensors (views of reused buffers) or None."""
        parts = []
        for k, b in enumerate(batches):
            if b is None:
                continue
            parts.append((k, self._stable_prep(k, b['prep']), self._buffers(k, b['prep']['nb'])))
        if not parts:
            raise ValueError('no rows to sweep')
        self._run(parts, FORWARD)
        lib = _lib.load()
        for k, prep, bufs in parts:
            sim, b = self.sims[k], batches[k]
            g = sim.geom
            _lib.check(lib.ss_residual_f32(sim._geom(prep['nb']), _lib.ptr(bufs['sim']), _lib.ptr(b['obs']),
                                           int(b.get('obs_repeat', 1)), _lib.ptr(b.get('v_mute')), g.dt, g.dx,
                                           t_mute, 0.03, _lib.ptr(bufs['res']), self._stream()), 'residual')
        self._run(parts, ADJOINT)
        out = [None] * len(batches)
        for k, prep, bufs in parts:
            sim = self.sims[k]
            g = sim.geom
            nb = prep['nb']
            grad = self._get(k, 'grad', (nb, g.ns, g.nz, g.nx))
            kk = float(np.float32((g.dt / g.dx) ** 2))
            _lib.check(lib.ss_gradient_f32(sim._geom(nb), _lib.ptr(bufs['gacc']), _lib.ptr(prep['v']),
                                           _lib.ptr(bufs['sg']), kk, _lib.ptr(grad), self._stream()), 'gradient')
            out[k] = (grad, bufs['illum'], bufs['sim'])
        return out

    def held_bytes(self):
        """Device bytes of the buffers this sweep keeps between calls."""
        return sum(t.numel() * t.element_size() for b in self._buf for t in b.values())

    def _round(self, k):
        """Propagations per round of group k's band launches on this device (0: the
        group runs on the tiled kernels, where batch size has no such granularity)."""
        if k not in self._rounds:
            n = _lib.load().ss_sweep_round_f32(ctypes.byref(self.sims[k]._geom(1)))
            self._rounds[k] = max(int(n), 0)
        return self._rounds[k]

    @staticmethod
    def _split(models, props, per_round, n, cap):
        """n model counts in [0, cap] summing to ``models``: the fewest band rounds
        (ceil(props * m / per_round) per chunk), then the most even; largest first."""
        if models == 0:
            return [0] * n
        if per_round <= 0:
            out = [min(cap, models - c * cap) for c in range(n)]
            return [max(m, 0) for m in out]
        best = {0: (0, 0, ())}
        for _ in range(n):
            nxt = {}
            for used, (r, sq, cs) in best.items():
                for m in range(0, min(cap, models - used) + 1):
                    cand = (r + -(-props * m // per_round), sq + m * m, cs + (m,))
                    if used + m not in nxt or cand[:2] < nxt[used + m][:2]:
                        nxt[used + m] = cand
            best = nxt
        return sorted(best[models][2], reverse=True)

    def _chunks(self, rows, repeats):
        """Split each group's rows into chunks of whole observation rows whose
        Laplacian histories fit the memory budget.  Each group's buffers are sized
        by its largest chunk, so the budget bounds the sum over groups of those.  A
        group whose propagations run in rounds (the band kernels, or the column-strip
        kernels' one workgroup per CU: ss_sweep_round_f32) gets chunk sizes that fill
        its rounds, and may take larger chunks than an even split where that saves
        rounds; the plan minimises the rounds (weighted by step count) over a few
        chunk counts.

        The plan is fixed per batch shape on first use, so later calls reuse the
        same buffers (and captured graphs)."""
        key = (tuple(rows), tuple(repeats))
        if key in self._plans:
            return self._plans[key]
        budget = self.max_history_bytes
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget = 0.6 * (free + self.held_bytes())
        groups = []  # (models, rows per model, propagations per model, per round, bytes per model, steps)
        for k, (nb, rep) in enumerate(zip(rows, repeats)):
            g = self.sims[k].geom
            per_model = rep * g.ns * g.nt * self.sims[k].H * self.sims[k].W * 4
            groups.append((nb // rep if nb else 0, rep, rep * g.ns, self._round(k) if nb else 0, per_model, g.nt))
        most = max([m for m, *_ in groups] + [1])
        n0 = max(1, math.ceil(self.history_bytes(rows) / budget))
        best, smallest = None, None
        for n in range(n0, most + 1):
            options = []  # per group: (weighted rounds, bytes held, counts), fewer rounds per larger cap
            for m, _, p, t, h, nt in groups:
                even = -(-m // n) if m else 0
                if t <= 0 or m == 0:
                    c = self._split(m, p, t, n, even)
                    options.append([(0, max(c) * h, c)])
                    continue
                opts, last = [], None
                for cap in range(even, m + 1):
                    c = self._split(m, p, t, n, cap)
                    r = sum(-(-p * x // t) for x in c)
                    if last is None or r < last:
                        opts.append((r * nt, max(c) * h, c))
                        last = r
                options.append(opts)
            for combo in itertools.product(*options):
                held = sum(o[1] for o in combo)
                plan = (sum(o[0] for o in combo), n, [o[2] for o in combo])
                if smallest is None or held < smallest[0]:
                    smallest = (held, plan)
                if held <= budget and (best is None or plan[0] < best[0]):
                    best = plan
            if best is not None and (n >= n0 + 3 or best[0] == 0):
                break
        _, n, counts = best if best is not None else smallest[1]
        spans = []
        for c, (_, rep, *_r) in zip(counts, groups):
            ends = np.cumsum([0] + list(c)) * rep
            spans.append([(int(ends[i]), int(ends[i + 1])) for i in range(n)])
        self._plans[key] = (n, spans)
        return n, spans

    @torch.no_grad()
    def images_device(self, batches, t_mute=0.12, eps=1e-4, laplacian=True):
        """RTM images of every group.

        ``batches[k]``: dict with ``v_bg`` ((nb, nz, nx) float64 velocities),
        ``obs`` (device (nb / obs_repeat, ns, nt, nx)), ``obs_repeat`` (rows per
        observation, default 1) and ``v_mute`` (device (nb,) or None), or None.
        Returns per group (img, illum) device tensors (nb, ns, nz, nx)."""
        if len(batches) != len(self.sims):
            raise ValueError('one batch (or None) per grid size is required')
        rows, reps = [], []
        for sim, b in zip(self.sims, batches):
            if b is None:
                rows.append(0)
                reps.append(1)
                continue
            rep = int(b.get('obs_repeat', 1))
            nb = len(b['v_bg'])
            g = sim.geom
            if rep < 1 or nb % rep or tuple(b['obs'].shape) != (nb // rep, g.ns, g.nt, g.nx):
                raise ValueError(f'observations must have shape {(nb // max(rep, 1), g.ns, g.nt, g.nx)}')
            rows.append(nb)
            reps.append(rep)
        nchunks, spans = self._chunks(rows, reps)
        imgs = [[] for _ in batches]
        illums = [[] for _ in batches]
        for c in range(nchunks):
            sub = []
            for k, b in enumerate(batches):
                a, e = spans[k][c] if rows[k] else (0, 0)
                if b is None or e <= a:
                    sub.append(None)
                    continue
                rep = reps[k]
                vm = b.get('v_mute')
                sub.append({'prep': self.sims[k]._prepare(np.asarray(b['v_bg'])[a:e]),
                            'obs': b['obs'][a // rep:e // rep], 'obs_repeat': rep,
                            'v_mute': None if vm is None else vm[a:e].contiguous()})
            res = self.gradients_device(sub, t_mute)
            for k, r in enumerate(res):
                if r is None:
                    continue
                grad, illum, _ = r
                img = self.sims[k]._image_device(grad, illum, eps, laplacian)
                imgs[k].append(img)
                illums[k].append(illum.clone())
        return [(torch.cat(i), torch.cat(l)) if i else None for i, l in zip(imgs, illums)]


def config_acquisitions():
    """The three grid sizes of the uneven-batch workload: 70 x 70 (nt 1000, 5 shots),
    100 x 140 (nt 1500, 7 shots) and 200 x 200 (nt 2000, 10 shots), dx 10 m, 15 Hz."""
    return [Acquisition(nz=70, nx=70, nt=1000, ns=5), Acquisition(nz=100, nx=140, nt=1500, ns=7),
            Acquisition(nz=200, nx=200, nt=2000, ns=10)]
