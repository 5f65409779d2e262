"""Batched finite-difference simulation and exact discrete-adjoint images on MI355X.

Drop-in for the reference module ``openfwi/physics/gpu.py``: the same names
(``Sim``, ``lap``, ``lap_adjoint``, ``fold_edge_padding``,
``imaging_memory_bytes``, ``calibrate``, ``resolve_device``, ``DEV``, ``SRC``,
``RICKER``, ``torch``), signatures, defaults, return types and error
messages.  The time loops run in hand-written gfx950 HIP kernels
(``libshotsweep.so``, called through ctypes on raw device pointers on the
current HIP stream); this module only forms the float64 coefficients on the
host exactly as the reference does, moves them to the device and sequences the
kernels.  There is no CPU path: a ``Sim`` on a non-GPU device raises.

Numerics: the kernels reproduce the reference's operation order, so FP32 and
FP64 gathers, gradients and images match the reference CPU path to the last
bit except for the reduction order of the padding fold.

Beyond the reference: ``Sim(..., geom=Acquisition(...))`` runs other grid
sizes, shot counts and step counts; ``history`` selects whether the adjoint
reads a stored Laplacian history ('full') or recomputes checkpointed segments
('checkpoint', the reference's scheme); 'auto' stores the full history when it
fits the device memory budget.
"""
import math

import numpy as np
import torch

from . import _lib
from .constants import DX, DT, NX, NZ, NT, F0, SRC_X, OPENFWI, Acquisition

DEV = torch.device('cuda' if torch.cuda.is_available() else 'cpu')
SRC = list(OPENFWI.src)
RICKER = OPENFWI.ricker()

__all__ = ['DEV', 'SRC', 'RICKER', 'Sim', 'lap', 'lap_adjoint', 'fold_edge_padding', 'imaging_memory_bytes',
           'calibrate', 'resolve_device', 'torch', 'DX', 'DT', 'NX', 'NZ', 'NT', 'F0', 'SRC_X']


def resolve_device(name='auto'):
    """Device selection with the reference's errors (gpu.py:20-30)."""
    if name == 'auto':
        return torch.device('cuda' if torch.cuda.is_available() else 'cpu')
    device = torch.device(name)
    if device.type == 'cuda' and not torch.cuda.is_available():
        raise RuntimeError('CUDA was requested, but this PyTorch installation has no available CUDA device')
    if device.type == 'mps':
        raise RuntimeError('MPS was requested, but no MPS device is available')
    if device.type not in ('cpu', 'cuda', 'mps'):
        raise ValueError(f'unsupported device: {name}')
    return device


def _require_gpu(t, what):
    if not t.is_cuda:
        raise RuntimeError(f'{what}: shotsweep_amd runs on a ROCm GPU; got a {t.device} tensor')


def _planes(p):
    """(..., H, W) tensor -> contiguous (n, H, W) view/copy and n."""
    p = p.contiguous()
    H, W = p.shape[-2:]
    n = p.numel() // (H * W) if p.numel() else 0
    return p, n, H, W


def lap(p):
    """Fourth-order Laplacian times dx^2, zero on the two outer cells (gpu.py:33-37)."""
    _require_gpu(p, 'lap')
    p, n, H, W = _planes(p)
    out = torch.empty_like(p)
    _lib.check(_lib.fn('lap', p.dtype)(n, H, W, _lib.ptr(p), _lib.ptr(out), _lib.stream_handle(p.device)), 'lap')
    return out


def lap_adjoint(p):
    """Transpose of lap, including its two-cell zero-output boundary (gpu.py:40-53)."""
    _require_gpu(p, 'lap_adjoint')
    p, n, H, W = _planes(p)
    out = torch.empty_like(p)
    _lib.check(_lib.fn('lap_adjoint', p.dtype)(n, H, W, _lib.ptr(p), _lib.ptr(out), _lib.stream_handle(p.device)),
               'lap_adjoint')
    return out


def fold_edge_padding(g, pad):
    """Transpose of np.pad(..., mode='edge') for the last two axes (gpu.py:56-67)."""
    _require_gpu(g, 'fold_edge_padding')
    g, n, H, W = _planes(g)
    nz, nx = H - 2 * pad, W - 2 * pad
    out = torch.empty(*g.shape[:-2], nz, nx, dtype=g.dtype, device=g.device)
    _lib.check(_lib.fn('fold_edge_padding', g.dtype)(n, nz, nx, pad, _lib.ptr(g), _lib.ptr(out),
                                                      _lib.stream_handle(g.device)), 'fold_edge_padding')
    return out


def imaging_memory_bytes(batch, pad=40, checkpoint_interval=40, bytes_per_value=4):
    """Conservative tensor-storage estimate of the reference scheme (gpu.py:70-73).

    Kept identical so that callers size batches as before; ``Sim.workspace_bytes``
    gives this implementation's own device footprint."""
    fields = 2 * math.ceil(NT / checkpoint_interval) + checkpoint_interval + 24
    return int(batch * 5 * bytes_per_value * ((NZ + 2 * pad) * (NX + 2 * pad) * fields + 3 * NT * NX))


class Sim:
    """Batched acoustic FD simulator, exact adjoint and RTM images (gpu.py:76-210)."""

    def __init__(self, wavelet=None, pad=40, R=1e-5, checkpoint_interval=40,
                 device=None, dtype=torch.float32, *, geom=None, history='auto', max_history_bytes=None,
                 impl=_lib.IMPL_AUTO, segment_impl='auto'):
        if pad < 2 or checkpoint_interval < 1:
            raise ValueError('pad must be >= 2 and checkpoint_interval must be positive')
        if history not in ('auto', 'full', 'checkpoint'):
            raise ValueError("history must be 'auto', 'full' or 'checkpoint'")
        if segment_impl not in ('auto', 'resident', 'tiled'):
            raise ValueError("segment_impl must be 'auto', 'resident' or 'tiled'")
        self.geom = OPENFWI if geom is None else geom
        self.pad = pad
        self.checkpoint_interval = checkpoint_interval
        self.device = DEV if device is None else resolve_device(str(device))
        if self.device.type != 'cuda':
            raise RuntimeError(f'shotsweep_amd.gpu.Sim runs its kernels on a ROCm GPU; device {self.device} '
                               'is not supported (there is no CPU path)')
        if dtype not in (torch.float32, torch.float64):
            raise ValueError('dtype must be torch.float32 or torch.float64')
        self.dtype = dtype
        self.history = history
        self.max_history_bytes = max_history_bytes
        self.impl = impl
        g = self.geom
        self.H, self.W = g.nz + 2 * pad, g.nx + 2 * pad
        A, B = self._damping(R)
        self.A = self._tensor(A)
        self.B = self._tensor(B)
        self.wav = np.asarray(g.ricker() if wavelet is None else wavelet)
        if self.wav.shape != (g.nt,) or not np.isfinite(self.wav).all() or not np.any(self.wav):
            raise ValueError(f'wavelet must contain {g.nt} finite samples and nonzero energy')
        self.w = self._tensor(self.wav * 1e3)
        self.si = torch.arange(g.ns, device=self.device)
        self.sx = torch.tensor(list(g.src), device=self.device) + pad
        self._ws = {}
        self._side = None    # side stream of the checkpointed adjoint (created on first use)
        self.segment_impl = segment_impl
        self._seg_tiles = {}  # direction -> (tile_h, tile_w, depth) of the tiled segment path
        self.events = None   # list -> (name, start, end) HIP events around each time-loop launch

    # ---------------------------------------------------------------- setup
    def _damping(self, R):
        """Quadratic sponge, float64 (gpu.py:85-90), one profile per axis."""
        pad, g = self.pad, self.geom

        def depth(n):
            i = np.arange(n)
            return np.clip(np.maximum(pad - i, i - (n - 1 - pad)), 0, None) / pad

        dz, dx = depth(self.H), depth(self.W)
        a = 3 * 4500 / (2 * pad * g.dx) * np.log(1 / R) * (dz[:, None] ** 2 + dx[None, :] ** 2) * g.dt / 2
        return 1 / (1 + a), 1 - a

    def _tensor(self, value):
        return torch.as_tensor(value, dtype=self.dtype, device=self.device)

    def _host(self, value):
        return torch.as_tensor(value, dtype=self.dtype)

    def _geom(self, nb):
        g = self.geom
        return _lib.make_geom(nb, g.ns, g.nz, g.nx, self.pad, g.nt, g.row, g.src)

    def _setup(self, v):
        """c2 (B,1,H,W) and amp (B,ns) (gpu.py:101-109)."""
        g = self.geom
        v = np.asarray(v, np.float64)
        if v.ndim != 3 or v.shape[1:] != (g.nz, g.nx) or not np.isfinite(v).all() or np.any(v <= 0):
            raise ValueError(f'velocities must be finite positive (B, {g.nz}, {g.nx}) arrays')
        p = self.pad
        # edge padding and vp^2 (dt/dx)^2 in float64 on the device (exact IEEE products, the
        # reference's order), then one rounding to the working dtype
        vd = torch.as_tensor(v, dtype=torch.float64, device=self.device)
        vp = torch.nn.functional.pad(vd[:, None], (p, p, p, p), mode='replicate')
        c2 = ((vp * vp) * ((g.dt / g.dx) ** 2)).to(self.dtype)
        amp = self._tensor(np.ascontiguousarray((v[:, g.row, list(g.src)] / 2000.0) ** 2))
        return c2, amp

    def _source_derivative(self, v):
        """2 v[row, src] / 2000^2 in the working dtype (gpu.py:168), formed on the host
        with the reference's CPU division."""
        g = self.geom
        sd = 2 * self._host(np.ascontiguousarray(np.asarray(v, np.float64)[:, g.row, list(g.src)])) / 2000.0 ** 2
        return sd.contiguous().to(self.device)

    def _scratch(self, name, shape, zero=False):
        t = self._ws.get(name)
        n = int(np.prod(shape))
        if t is None or t.numel() < n:
            self._ws.pop(name, None)
            t = torch.empty(n, dtype=self.dtype, device=self.device)
            self._ws[name] = t
        out = t[:n].view(*shape)
        if zero:
            out.zero_()
        return out

    def _side_stream(self, main):
        """Stream of the segment recomputes: a second stream beside the caller's, or the
        caller's own while HIP events are recorded (each event interval then spans one
        launch alone)."""
        if self.events is not None:
            return main
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.device)  # default priority
        return self._side

    def _stable(self, name, t):
        """t's values in a cached workspace of the same shape.  The sweep entry points
        replay hipGraphs keyed by buffer addresses, so tensors made per call are
        copied into buffers that keep their address from call to call."""
        buf = self._scratch(name, tuple(t.shape))
        if buf.data_ptr() != t.data_ptr():
            buf.copy_(t)
        return buf

    def _stable_prep(self, prep):
        return dict(prep, c2=self._stable('seg_c2', prep['c2']), amp=self._stable('seg_amp', prep['amp']),
                    sd=self._stable('seg_sd', prep['sd']))

    def release(self):
        """Free cached device workspaces."""
        self._ws.clear()

    # ------------------------------------------------------------- kernels
    def _stream(self):
        return _lib.stream_handle(self.device)

    def _timed(self, name):
        if self.events is None:
            return None
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        self.events.append((name, start, end))
        return end

    def _forward(self, geom, t0, t1, c2, amp, p0, p1, gathers=None, hist=None, ckpt=None, every=1, impl=None):
        f = _lib.fn('forward', self.dtype)
        end = self._timed('forward')
        _lib.check(f(geom, t0, t1, _lib.ptr(c2), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w),
                     _lib.ptr(amp), _lib.ptr(p0), _lib.ptr(p1), _lib.ptr(gathers), _lib.ptr(hist), _lib.ptr(ckpt),
                     every, self.impl if impl is None else impl, self._stream()), 'forward')
        if end is not None:
            end.record()

    def _adjoint(self, geom, t0, t1, c2, sd, res, hist, m1, m2, gacc, illum, sg, impl=None):
        f = _lib.fn('adjoint', self.dtype)
        end = self._timed('adjoint')
        _lib.check(f(geom, t0, t1, _lib.ptr(c2), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w),
                     _lib.ptr(sd), _lib.ptr(res), _lib.ptr(hist), _lib.ptr(m1), _lib.ptr(m2), _lib.ptr(gacc),
                     _lib.ptr(illum), _lib.ptr(sg), self.impl if impl is None else impl, self._stream()), 'adjoint')
        if end is not None:
            end.record()

    def _step(self, p0, p1, c2, amp, it):
        """One damped leapfrog step; returns (p2, L) (gpu.py:111-115)."""
        nb = p1.shape[0]
        a, b = p0.contiguous().clone(), p1.contiguous().clone()
        L = torch.empty(nb, self.geom.ns, 1, self.H, self.W, dtype=self.dtype, device=self.device)
        self._forward(self._geom(nb), it, it + 1, c2.contiguous(), amp.contiguous(), a, b, hist=L,
                      impl=_lib.IMPL_STEPWISE)
        return b, L[:, :, 0]

    def forward(self, v, store=False):
        """Return gathers, optional wavefield checkpoints, and squared Courant numbers.

        Checkpoints hold the states (p0, p1) before every checkpoint_interval-th
        step, as views into one device buffer (gpu.py:117-136)."""
        g = self.geom
        c2, amp = self._setup(v)
        nb = len(c2)
        shape = (nb, g.ns, self.H, self.W)
        p0 = torch.zeros(shape, dtype=self.dtype, device=self.device)
        p1 = torch.zeros_like(p0)
        out = torch.empty(nb, g.ns, g.nt, g.nx, dtype=self.dtype, device=self.device)
        ci = self.checkpoint_interval
        if store:
            nck = -(-g.nt // ci)
            ck = torch.empty(nb, g.ns, nck, 2, self.H, self.W, dtype=self.dtype, device=self.device)
            self._forward(self._geom(nb), 0, g.nt, c2, amp, p0, p1, gathers=out, ckpt=ck, every=ci)
            checkpoints = {k * ci: (ck[:, :, k, 0], ck[:, :, k, 1]) for k in range(nck)}
        else:
            self._forward(self._geom(nb), 0, g.nt, c2, amp, p0, p1, gathers=out)
            checkpoints = None
        return out, checkpoints, c2

    @torch.no_grad()
    def gathers(self, v, chunk=20):
        return np.concatenate([self.forward(v[i:i + chunk])[0].cpu().numpy() for i in range(0, len(v), chunk)])

    def _check_mute(self, v_mute):
        v_mute = np.asarray(v_mute)
        if not np.isfinite(v_mute).all() or np.any(v_mute <= 0):
            raise ValueError('mute velocities must be finite and positive')
        return self._tensor(v_mute).reshape(-1).contiguous()

    def mute_weights(self, v_mute, t_mute=0.12):
        """Direct-wave cosine mute, shape (len(v_mute), ns, nt, nx) (gpu.py:142-150)."""
        g = self.geom
        vm = self._check_mute(v_mute)
        out = torch.empty(len(vm), g.ns, g.nt, g.nx, dtype=self.dtype, device=self.device)
        f = _lib.fn('mute_weights', self.dtype)
        _lib.check(f(self._geom(len(vm)), _lib.ptr(vm), g.dt, g.dx, t_mute, 0.03, _lib.ptr(out), self._stream()),
                   'mute_weights')
        return out

    # --------------------------------------------------------- memory plan
    def workspace_bytes(self, nb, history=None):
        """Device bytes of one gradient/images call on nb rows under a history mode."""
        itemsize = torch.empty((), dtype=self.dtype).element_size()
        return _lib.workspace_bytes(self._geom(nb), self._segment(nb, history), itemsize)

    def _segment(self, nb, history=None):
        """Steps per adjoint segment: nt (whole history stored) or the checkpoint interval."""
        g = self.geom
        mode = self.history if history is None else history
        if mode == 'checkpoint':
            return min(self.checkpoint_interval, g.nt)
        if mode == 'full':
            return g.nt
        itemsize = torch.empty((), dtype=self.dtype).element_size()
        full = nb * g.ns * self.H * self.W * itemsize * g.nt
        budget = self.max_history_bytes
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget = 0.6 * free
        return g.nt if full <= budget else min(self.checkpoint_interval, g.nt)

    # ----------------------------------------------------------- gradients
    def _prepare(self, v_bg):
        c2, amp = self._setup(v_bg)
        return {'c2': c2, 'amp': amp, 'sd': self._source_derivative(v_bg),
                'v': self._tensor(np.asarray(v_bg, np.float64)).contiguous(), 'nb': len(c2)}

    def _observed(self, obs, nb):
        g = self.geom
        observed = self._tensor(np.asarray(obs))
        if observed.shape != (nb, g.ns, g.nt, g.nx) or not torch.isfinite(observed).all():
            raise ValueError(f'observations must be finite with shape {(nb, g.ns, g.nt, g.nx)}')
        return observed.contiguous()

    def _forward_tape(self, prep, segment=None, fresh=False):
        """Forward pass that keeps what the adjoint needs.

        Returns (gathers, tape): tape holds either the whole Laplacian history
        (one adjoint launch) or checkpoints every ``segment`` steps (the
        adjoint recomputes each segment, gpu.py:169-176).  ``fresh`` puts the
        gathers and the tape in new tensors instead of the cached workspaces
        (for callers that keep them, e.g. autograd)."""
        g = self.geom
        nb, c2, amp = prep['nb'], prep['c2'], prep['amp']
        geom = self._geom(nb)
        H, W = self.H, self.W
        seg = self._segment(nb) if segment is None else min(segment, g.nt)
        new = (lambda name, shape, zero=False: torch.zeros(shape, dtype=self.dtype, device=self.device)) if fresh \
            else self._scratch
        p0 = self._scratch('p0', (nb, g.ns, H, W), zero=True)
        p1 = self._scratch('p1', (nb, g.ns, H, W), zero=True)
        sim = new('sim', (nb, g.ns, g.nt, g.nx))
        if seg >= g.nt:
            hist = new('hist', (nb, g.ns, g.nt, H, W))
            self._forward(geom, 0, g.nt, c2, amp, p0, p1, gathers=sim, hist=hist)
            return sim, {'seg': g.nt, 'hist': hist}
        nck = -(-g.nt // seg)
        ck = new('ckpt', (nb, g.ns, nck, 2, H, W))
        if self._tiled_segments(nb):
            self._forward_tape_tiled(prep, seg, sim, ck)
        else:
            self._forward(geom, 0, g.nt, c2, amp, p0, p1, gathers=sim, ckpt=ck, every=seg)
        return sim, {'seg': seg, 'ckpt': ck}

    def _adjoint_tape(self, prep, tape, res):
        """Exact adjoint of the taped forward for the data-space seed ``res``
        (nb, ns, nt, nx), imaging condition and fold (gpu.py:161-188).
        Returns (grad (nb, ns, nz, nx), illum (nb, ns, nz, nx))."""
        g = self.geom
        nb, c2, amp, sd = prep['nb'], prep['c2'], prep['amp'], prep['sd']
        geom = self._geom(nb)
        H, W = self.H, self.W
        m1 = self._scratch('m1', (nb, g.ns, H, W), zero=True)
        m2 = self._scratch('m2', (nb, g.ns, H, W), zero=True)
        gacc = self._scratch('gacc', (nb, g.ns, H, W), zero=True)
        illum = torch.zeros(nb, g.ns, g.nz, g.nx, dtype=self.dtype, device=self.device)
        sg = torch.zeros(nb, g.ns, dtype=self.dtype, device=self.device)
        seg = tape['seg']
        if 'hist' in tape:
            self._adjoint(geom, 0, g.nt, c2, sd, res, tape['hist'], m1, m2, gacc, illum, sg)
        elif self._tiled_segments(nb):
            self._segments_tiled(prep, tape, res, gacc, illum, sg)
        else:
            self._segments(geom, prep, tape, res, m1, m2, gacc, illum, sg)
        grad = torch.empty(nb, g.ns, g.nz, g.nx, dtype=self.dtype, device=self.device)
        k = np.dtype(np.float32 if self.dtype == torch.float32 else np.float64).type((g.dt / g.dx) ** 2)
        f = _lib.fn('gradient', self.dtype)
        _lib.check(f(geom, _lib.ptr(gacc), _lib.ptr(prep['v']), _lib.ptr(sg), float(k), _lib.ptr(grad),
                     self._stream()), 'gradient')
        return grad, illum

    def _segments(self, geom, prep, tape, res, m1, m2, gacc, illum, sg):
        """Checkpointed adjoint (gpu.py:169-184): for each segment, last first,
        recompute its forward from the checkpoint, keeping the Laplacians, then
        run the adjoint over it.  The recompute of segment k-1 runs on a side
        stream while the adjoint of segment k runs on the caller's stream (they
        touch different buffers; two segment buffers alternate), so a batch
        that leaves CUs idle overlaps the two.  Each launch computes exactly
        what it computes in sequence, so results do not depend on the overlap."""
        g = self.geom
        nb, c2, amp, sd = prep['nb'], prep['c2'], prep['amp'], prep['sd']
        H, W = self.H, self.W
        ck, seg = tape['ckpt'], tape['seg']
        nseg = ck.shape[2]
        main = torch.cuda.current_stream(self.device)
        side = self._side_stream(main)
        bufs = [(self._scratch(f'p0_{i}', (nb, g.ns, H, W)), self._scratch(f'p1_{i}', (nb, g.ns, H, W)),
                 self._scratch(f'hist_seg{i}', (nb, g.ns, seg, H, W))) for i in (0, 1)]
        freed = [None, None]        # event after the adjoint that last read segment buffer i
        ready = {}

        def span(k):
            return k * seg, min(g.nt, (k + 1) * seg)

        def view(i, k):
            t0, t1 = span(k)
            return bufs[i][2].reshape(-1)[:nb * g.ns * (t1 - t0) * H * W].view(nb, g.ns, t1 - t0, H, W)

        def recompute(k):
            i = k % 2
            t0, t1 = span(k)
            if freed[i] is not None:
                side.wait_event(freed[i])
            with torch.cuda.stream(side):
                p0, p1, _ = bufs[i]
                p0.copy_(ck[:, :, k, 0])
                p1.copy_(ck[:, :, k, 1])
                self._forward(geom, t0, t1, c2, amp, p0, p1, hist=view(i, k))
                ready[k] = torch.cuda.Event()
                ready[k].record(side)

        side.wait_stream(main)      # checkpoints, residual and zeroed state are ready
        recompute(nseg - 1)
        for k in reversed(range(nseg)):
            if k > 0:
                recompute(k - 1)
            main.wait_event(ready.pop(k))
            t0, t1 = span(k)
            self._adjoint(geom, t0, t1, c2, sd, res, view(k % 2, k), m1, m2, gacc, illum, sg)
            freed[k % 2] = torch.cuda.Event()
            freed[k % 2].record(main)
        main.wait_stream(side)      # later allocations on the caller's stream come after the side work

    def _tiled_segments(self, nb):
        """Whether checkpoint segments run on the tiled kernels: FP32 on an even-sized
        grid, and (for 'auto') a batch of at most half as many propagations as the
        GPU has CUs, which the one-workgroup-per-propagation kernels would leave
        mostly idle."""
        g = self.geom
        fits = (self.dtype == torch.float32 and self.H * self.W <= 131072 and self.pad <= 44
                and g.nt * self.H * self.W < 2 ** 29)          # the tiled kernels' limits
        if self.segment_impl == 'tiled':
            if not fits:
                raise ValueError('tiled segments need FP32 and H*W <= 131072, pad <= 44, nt*H*W < 2^29')
            return True
        if self.segment_impl == 'resident' or not fits or self.impl != _lib.IMPL_AUTO:
            return False
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return 2 * nb * g.ns <= cus

    def _sweep_group(self, prep, direction, t0, t1, resume, phase, state, hist, res=None, gacc=None, illum=None,
                     sg=None, gathers=None):
        grp = (_lib.SSSweepGroup * 1)()
        g = grp[0]
        g.geom = self._geom(prep['nb'])
        g.tile_h, g.tile_w, g.depth = self._seg_tiles.get(direction, (0, 0, 0))
        g.c2, g.A, g.B, g.w = _lib.ptr(prep['c2']), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w)
        g.amp = _lib.ptr(prep['amp'] if direction == 0 else prep['sd'])
        g.state, g.hist = _lib.ptr(state), _lib.ptr(hist)
        if direction == 1:
            g.res, g.gacc, g.illum, g.src_grad = _lib.ptr(res), _lib.ptr(gacc), _lib.ptr(illum), _lib.ptr(sg)
        else:
            g.gathers = _lib.ptr(gathers)
        g.t0, g.t1, g.resume, g.phase = t0, t1, int(resume), phase
        return grp

    def _seg_tune(self, prep, direction, t0, t1, state, hist, res=None, gacc=None, illum=None, sg=None):
        """Tile shape and depth of the segment sweeps, tuned once per Sim and direction.
        The trial launches overwrite the state buffers and the adjoint outputs."""
        if direction not in self._seg_tiles:
            grp = self._sweep_group(prep, direction, t0, t1, 0, 0, state, hist, res, gacc, illum, sg)
            _lib.check(_lib.load().ss_sweep_autotune_f32(1, grp, direction, 0, self._stream()), 'sweep autotune')
            self._seg_tiles[direction] = (grp[0].tile_h, grp[0].tile_w, grp[0].depth)

    def _forward_tape_tiled(self, prep, seg, sim, ck):
        """Checkpointed forward (gpu.py:126-136 with checkpoints) on the tiled kernels,
        for batches too small to fill the GPU one workgroup per propagation: per
        segment, ss_sweep_forward_f32 advances [t0, t1) from the previous segment's
        state without keeping Laplacians and records the gathers; the state at each
        segment start is copied out as that segment's checkpoint.  The gathers and
        checkpoints equal the resident forward's bit for bit."""
        g = self.geom
        nb = prep['nb']
        H, W = self.H, self.W
        plane = nb * g.ns * H * W
        prep = self._stable_prep(prep)
        gathers = self._scratch('seg_gathers', tuple(sim.shape))
        fst = self._scratch('seg_fstate', (4 * plane,))
        states = fst.view(2, 2, nb, g.ns, H, W)        # buffer, (p0, p1), rows, shots
        lib = _lib.load()
        st = self._stream()
        # trials overwrite the state buffers: tune before the first segment
        self._seg_tune(prep, 0, 0, min(seg, g.nt), fst, self._scratch('hist_seg0', (nb, g.ns, seg, H, W)))
        depth = self._seg_tiles[0][2]
        ck[:, :, 0].zero_()
        phase = 0
        for k in range(ck.shape[2]):
            t0, t1 = k * seg, min(g.nt, (k + 1) * seg)
            if k > 0:
                ck[:, :, k].copy_(states[phase].permute(1, 2, 0, 3, 4))
            end = self._timed('forward')
            _lib.check(lib.ss_sweep_forward_f32(1, self._sweep_group(prep, 0, t0, t1, k > 0, phase, fst, None,
                                                                     gathers=gathers), 1, st), 'sweep forward')
            if end is not None:
                end.record()
            phase = (phase - (-(t1 - t0) // depth)) & 1
        if gathers.data_ptr() != sim.data_ptr():
            sim.copy_(gathers)

    def _segments_tiled(self, prep, tape, res, gacc, illum, sg):
        """Checkpointed adjoint on the tiled kernels (gpu.py:169-184), for batches too
        small to fill the GPU one workgroup per propagation: per segment, last
        first, ss_sweep_forward_f32 recomputes [t0, t1) from the checkpoint
        (resume) and keeps the Laplacians, then ss_sweep_adjoint_f32 runs [t0, t1)
        backwards from the previous segment's adjoint state, every propagation
        split into tiles over all CUs.  As in the resident loop, the recompute of
        segment k-1 runs on a side stream beside the adjoint of segment k (two
        alternating history buffers), which fills the gaps between the short
        dependent launches of each.  Per cell the arithmetic is that of the
        resident path, so the results are the same bit for bit."""
        g = self.geom
        nb = prep['nb']
        H, W = self.H, self.W
        ck, seg = tape['ckpt'], tape['seg']
        nseg = ck.shape[2]
        plane = nb * g.ns * H * W
        prep = self._stable_prep(prep)
        res = self._stable('seg_res', res)
        illum_out, sg_out = illum, sg
        illum = self._scratch('seg_illum', tuple(illum.shape))   # the first adjoint call overwrites them
        sg = self._scratch('seg_sg', tuple(sg.shape))
        fst = self._scratch('seg_fstate', (4 * plane,))     # forward: 2 buffers x (p0, p1)
        ast = self._scratch('seg_astate', (4 * plane,))     # adjoint: 2 buffers x (m1, m2)
        hists = [self._scratch(f'hist_seg{i}', (nb, g.ns, seg, H, W)) for i in (0, 1)]
        lib = _lib.load()
        main = torch.cuda.current_stream(self.device)
        side = self._side_stream(main)
        freed = [None, None]        # event after the adjoint that last read history buffer i
        ready = {}

        def span(k):
            return k * seg, min(g.nt, (k + 1) * seg)

        def view(k):
            t0, t1 = span(k)
            return hists[k % 2].reshape(-1)[:nb * g.ns * (t1 - t0) * H * W]

        def recompute(k):
            t0, t1 = span(k)
            if freed[k % 2] is not None:
                side.wait_event(freed[k % 2])
            with torch.cuda.stream(side):
                fst[:2 * plane].view(2, nb, g.ns, H, W).copy_(ck[:, :, k].permute(2, 0, 1, 3, 4))
                end = self._timed('forward')
                _lib.check(lib.ss_sweep_forward_f32(1, self._sweep_group(prep, 0, t0, t1, 1, 0, fst, view(k)), 1,
                                                    self._stream()), 'sweep forward')
                if end is not None:
                    end.record()
                ready[k] = torch.cuda.Event()
                ready[k].record(side)

        self._seg_tune(prep, 0, *span(nseg - 1), fst, hists[0])   # forward trials overwrite fst: tune first
        side.wait_stream(main)      # checkpoints, seed and zeroed state are ready
        recompute(nseg - 1)
        if 1 not in self._seg_tiles:
            # adjoint trials read the last segment's history; the first real adjoint call
            # (resume 0) then overwrites the outputs the trials left
            main.wait_event(ready[nseg - 1])
            self._seg_tune(prep, 1, *span(nseg - 1), ast, view(nseg - 1), res, gacc, illum, sg)
        phase = 0                   # adjoint state buffer that holds (m1, m2)
        for k in reversed(range(nseg)):
            if k > 0:
                recompute(k - 1)
            main.wait_event(ready.pop(k))
            t0, t1 = span(k)
            end = self._timed('adjoint')
            _lib.check(lib.ss_sweep_adjoint_f32(1, self._sweep_group(prep, 1, t0, t1, k < nseg - 1, phase, ast,
                                                                    view(k), res, gacc, illum, sg), 1,
                                                self._stream()), 'sweep adjoint')
            if end is not None:
                end.record()
            freed[k % 2] = torch.cuda.Event()
            freed[k % 2].record(main)
            phase = (phase - (-(t1 - t0) // self._seg_tiles[1][2])) & 1
        main.wait_stream(side)      # later allocations on the caller's stream come after the side work
        illum_out.copy_(illum)
        sg_out.copy_(sg)

    def _gradient_device(self, prep, observed=None, obs_repeat=1, v_mute=None, t_mute=0.12, res=None):
        """Forward, muted residual, exact adjoint and fold on the device.

        Either ``observed`` (device, (nb/obs_repeat, ns, nt, nx)) or a ready
        data-space seed ``res`` (nb, ns, nt, nx) is given.  Returns
        (grad (nb,ns,nz,nx), illum (nb,ns,nz,nx), sim gathers)."""
        g = self.geom
        nb = prep['nb']
        sim, tape = self._forward_tape(prep)
        if res is None:
            res = self._scratch('res', (nb, g.ns, g.nt, g.nx))
            f = _lib.fn('residual', self.dtype)
            _lib.check(f(self._geom(nb), _lib.ptr(sim), _lib.ptr(observed), obs_repeat, _lib.ptr(v_mute), g.dt,
                         g.dx, t_mute, 0.03, _lib.ptr(res), self._stream()), 'residual')
        grad, illum = self._adjoint_tape(prep, tape, res)
        return grad, illum, sim

    def _gradient(self, v_bg, obs, v_mute=None, t_mute=0.12):
        """Per-shot misfit gradient and illumination as device tensors (gpu.py:152-188)."""
        prep = self._prepare(v_bg)
        observed = self._observed(obs, prep['nb'])
        vm = None if v_mute is None else self._check_mute(v_mute)
        grad, illum, _ = self._gradient_device(prep, observed, 1, vm, t_mute)
        return grad, illum

    @torch.no_grad()
    def gradient(self, v_bg, obs, v_mute=None, t_mute=0.12):
        """Per-shot derivative of .5*sum(weight*(sim-obs)**2), with fixed mute weights."""
        grad, illum = self._gradient(v_bg, obs, v_mute, t_mute)
        return grad.cpu().numpy(), illum.cpu().numpy()

    def _image_device(self, grad, illum, eps=1e-4, laplacian=True):
        nb = grad.shape[0]
        img = torch.empty_like(grad)
        f = _lib.fn('image', self.dtype)
        e = float(np.dtype(np.float32 if self.dtype == torch.float32 else np.float64).type(eps))
        _lib.check(f(self._geom(nb), _lib.ptr(grad), _lib.ptr(illum), e, int(bool(laplacian)), _lib.ptr(img),
                     self._stream()), 'image')
        return img

    @torch.no_grad()
    def images(self, v_bg, obs, v_mute=None, t_mute=0.12, eps=1e-4, laplacian=True):
        """Return illumination-normalized negative gradients and source illumination.

        The direct-wave cosine mute has a 30 ms ramp starting at offset/v_mute
        + t_mute. Its weights are fixed during differentiation.
        """
        if eps <= 0:
            raise ValueError('illumination regularization eps must be positive')
        grad, illum = self._gradient(v_bg, obs, v_mute, t_mute)
        img = self._image_device(grad, illum, eps, laplacian)
        return img.cpu().numpy(), illum.cpu().numpy()

    @torch.no_grad()
    def images_device(self, prep, observed, obs_repeat=1, v_mute=None, t_mute=0.12, eps=1e-4, laplacian=True):
        """Device-resident images: ``prep`` from ``_prepare``, ``observed`` a device tensor
        (nb/obs_repeat, ns, nt, nx) reused by obs_repeat consecutive rows, ``v_mute`` a
        device tensor (nb,).  Returns (img, illum) device tensors."""
        grad, illum, _ = self._gradient_device(prep, observed, obs_repeat, v_mute, t_mute)
        return self._image_device(grad, illum, eps, laplacian), illum

    @torch.no_grad()
    def born(self, v, dv):
        """Tangent-linear gathers d(gathers)/dv . dv (no reference counterpart)."""
        g = self.geom
        c2, amp = self._setup(v)
        vv, dvv = np.asarray(v, np.float64), np.asarray(dv, np.float64)
        p = self.pad
        vp = np.pad(vv, ((0, 0), (p, p), (p, p)), mode='edge')
        dvp = np.pad(dvv, ((0, 0), (p, p), (p, p)), mode='edge')
        dc2 = self._tensor(2 * vp * dvp * (g.dt / g.dx) ** 2)
        damp = self._tensor(np.ascontiguousarray(2 * vv[:, g.row, list(g.src)] * dvv[:, g.row, list(g.src)]
                                                 / 2000.0 ** 2))
        nb = len(c2)
        shape = (nb, g.ns, self.H, self.W)
        p0, p1, q0, q1 = (torch.zeros(shape, dtype=self.dtype, device=self.device) for _ in range(4))
        out = torch.empty(nb, g.ns, g.nt, g.nx, dtype=self.dtype, device=self.device)
        f = _lib.fn('born', self.dtype)
        _lib.check(f(self._geom(nb), 0, g.nt, _lib.ptr(c2), _lib.ptr(dc2), _lib.ptr(self.A), _lib.ptr(self.B),
                     _lib.ptr(self.w), _lib.ptr(amp), _lib.ptr(damp), _lib.ptr(p0), _lib.ptr(p1), _lib.ptr(q0),
                     _lib.ptr(q1), _lib.ptr(out), self._stream()), 'born')
        return out


def calibrate(models, gathers, eps=1e-3):
    """One effective source wavelet pooled over training models with known velocity (gpu.py:213-221)."""
    s = Sim(RICKER)
    sims = s.gathers(np.asarray(models, np.float64))
    Sf = np.fft.rfft(sims, axis=2)
    Of = np.fft.rfft(np.asarray(gathers, np.float64), axis=2)
    num = (np.conj(Sf) * Of).sum(axis=(0, 1, 3))
    den = (np.abs(Sf) ** 2).sum(axis=(0, 1, 3))
    return np.fft.irfft(np.fft.rfft(RICKER) * num / (den + eps * den.max()), n=NT)
