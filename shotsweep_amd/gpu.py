"""Batched finite-difference simulation and exact discrete-adjoint images on MI355X.

Drop-in for the reference module ``openfwi/physics/gpu.py``: the same names
(``Sim``, ``lap``, ``lap_adjoint``, ``fold_edge_padding``,
``imaging_memory_bytes``, ``calibrate``, ``resolve_device``, ``DEV``, ``SRC``,
``RICKER``, ``torch``), signatures, defaults, return types and error
messages.  The time loops run in hand-written gfx950 HIP kernels
(``libshotsweep.so``, called through ctypes on raw device pointers on the
current HIP stream); this module only forms the float64 coefficients on the
host exactly as the reference does, moves them to the device and sequences the
kernels.  There is no CPU path: ``Sim(device='cpu')`` and CPU tensors handed to
``lap``, ``lap_adjoint`` and ``fold_edge_padding`` run the same GPU kernels and
return on the CPU; without a GPU they raise.

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
    if device.type == 'mps' and not torch.backends.mps.is_available():
        raise RuntimeError('MPS was requested, but no MPS device is available')
    if device.type not in ('cpu', 'cuda', 'mps'):
        raise ValueError(f'unsupported device: {name}')
    return device


def _kernel_device(device):
    """Device the kernels run on for a requested device: a ROCm device itself, or for
    'cpu' the current ROCm device (inputs are staged there and results returned on
    the requested device).  There is no CPU path: without a GPU this raises."""
    device = torch.device(device)
    if device.type == 'cuda':
        return device
    if not torch.cuda.is_available():
        raise RuntimeError(f'shotsweep_amd runs its kernels on a ROCm GPU; device {device} was requested and no GPU '
                           'is available (there is no CPU path)')
    return torch.device('cuda', torch.cuda.current_device())


def _planes(p):
    """(..., H, W) tensor -> contiguous (n, H, W) view/copy and n."""
    p = p.contiguous()
    H, W = p.shape[-2:]
    n = p.numel() // (H * W) if p.numel() else 0
    return p, n, H, W


def _on_kernel_device(t):
    """t on the device the kernels run on (a copy for CPU tensors)."""
    return t if t.is_cuda else t.to(_kernel_device(t.device))


def lap(p):
    """Fourth-order Laplacian times dx^2, zero on the two outer cells (gpu.py:33-37).
    CPU tensors are staged through the GPU kernel and returned on the CPU."""
    src = p
    p, n, H, W = _planes(_on_kernel_device(p))
    out = torch.empty_like(p)
    _lib.check(_lib.fn('lap', p.dtype)(n, H, W, _lib.ptr(p), _lib.ptr(out), _lib.stream_handle(p.device)), 'lap')
    return out.to(src.device)


def lap_adjoint(p):
    """Transpose of lap, including its two-cell zero-output boundary (gpu.py:40-53).
    CPU tensors are staged through the GPU kernel and returned on the CPU."""
    src = p
    p, n, H, W = _planes(_on_kernel_device(p))
    out = torch.empty_like(p)
    _lib.check(_lib.fn('lap_adjoint', p.dtype)(n, H, W, _lib.ptr(p), _lib.ptr(out), _lib.stream_handle(p.device)),
               'lap_adjoint')
    return out.to(src.device)


def fold_edge_padding(g, pad):
    """Transpose of np.pad(..., mode='edge') for the last two axes (gpu.py:56-67).
    CPU tensors are staged through the GPU kernel and returned on the CPU."""
    src = g
    g, n, H, W = _planes(_on_kernel_device(g))
    nz, nx = H - 2 * pad, W - 2 * pad
    out = torch.empty(*g.shape[:-2], nz, nx, dtype=g.dtype, device=g.device)
    _lib.check(_lib.fn('fold_edge_padding', g.dtype)(n, nz, nx, pad, _lib.ptr(g), _lib.ptr(out),
                                                      _lib.stream_handle(g.device)), 'fold_edge_padding')
    return out.to(src.device)


def imaging_memory_bytes(batch, pad=40, checkpoint_interval=40, bytes_per_value=4):
    """Conservative tensor-storage estimate of the reference scheme (gpu.py:70-73).

    Kept identical so that callers size batches as before; ``Sim.workspace_bytes``
    gives this implementation's own device footprint."""
    fields = 2 * math.ceil(NT / checkpoint_interval) + checkpoint_interval + 24
    return int(batch * 5 * bytes_per_value * ((NZ + 2 * pad) * (NX + 2 * pad) * fields + 3 * NT * NX))


class Sim:
    """Batched acoustic FD simulator, exact adjoint and RTM images (gpu.py:76-210).

    ``device`` is where results are returned, as in the reference; the kernels run
    on that device when it is a ROCm GPU and on the current ROCm GPU for 'cpu'
    (inputs are staged to it, results come back to the CPU)."""

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
        self.kdev = _kernel_device(self.device)
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
        self.si = torch.arange(g.ns, device=self.kdev)
        self.sx = torch.tensor(list(g.src), device=self.kdev) + pad
        self._ws = {}
        self.segment_impl = segment_impl
        self.timing = False      # True: HIP events around every time-loop launch of the native calls
        self.timing_log = []     # per such call: {'forward': (launches, ms), 'adjoint': (launches, ms)}

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
        return torch.as_tensor(value, dtype=self.dtype, device=self.kdev)

    def _out(self, t):
        """A result on the requested device."""
        return t if t.device == self.device else t.to(self.device)

    @property
    def _itemsize(self):
        return 4 if self.dtype == torch.float32 else 8

    def _geom(self, nb):
        g = self.geom
        return _lib.make_geom(nb, g.ns, g.nz, g.nx, self.pad, g.nt, g.row, g.src)

    def _velocity(self, v):
        """Velocities (B, nz, nx) as a float64 tensor on the kernel device, validated as
        the reference does (gpu.py:102-104), and the float64 host array when the input
        came from the host (None for device tensors, which stay on the device)."""
        g = self.geom
        host = None
        if torch.is_tensor(v) and v.is_cuda:
            vd = v.detach().to(self.kdev, torch.float64)
            bad = vd.dim() != 3 or tuple(vd.shape[1:]) != (g.nz, g.nx) or \
                bool((~torch.isfinite(vd) | (vd <= 0)).any())
        else:
            host = np.asarray(v.detach().cpu() if torch.is_tensor(v) else v, np.float64)
            bad = host.ndim != 3 or host.shape[1:] != (g.nz, g.nx) or not np.isfinite(host).all() or \
                np.any(host <= 0)
            vd = None if bad else torch.as_tensor(host, device=self.kdev)
        if bad:
            raise ValueError(f'velocities must be finite positive (B, {g.nz}, {g.nx}) arrays')
        return vd, host

    def _coefficients(self, vd, host=None):
        """c2 (B,H,W) and amp (B,ns) (gpu.py:105-108).  c2: edge padding and
        vp*vp*(dt/dx)^2 in float64 on the device (exact IEEE products, the reference's
        order), rounded once.  amp = (v/2000)^2 in float64, rounded once: on the host
        with NumPy as the reference for host input; for device input with a
        tensor-by-tensor division (torch turns a division by a Python scalar into a
        reciprocal product, which can differ in the last bit)."""
        g, p = self.geom, self.pad
        vp = torch.nn.functional.pad(vd[:, None], (p, p, p, p), mode='replicate')[:, 0]
        c2 = ((vp * vp) * ((g.dt / g.dx) ** 2)).to(self.dtype).contiguous()
        if host is not None:
            amp = self._tensor(np.ascontiguousarray((host[:, g.row, list(g.src)] / 2000.0) ** 2))
        else:
            vs = vd[:, g.row][:, list(g.src)]
            a = vs / torch.full_like(vs, 2000.0)
            amp = (a * a).to(self.dtype).contiguous()
        return c2, amp

    def _setup(self, v):
        """c2 (B,H,W) and amp (B,ns) (gpu.py:101-109)."""
        return self._coefficients(*self._velocity(v))

    def _source_derivative(self, vd, host=None):
        """2 v[row, src] / 2000^2 in the working dtype (gpu.py:168): v rounded to the
        dtype, doubled (exact) and divided by 2000^2, on the host with the reference's
        CPU division for host input, else by a tensor-by-tensor division."""
        g = self.geom
        if host is not None:
            vs = torch.as_tensor(np.ascontiguousarray(host[:, g.row, list(g.src)]), dtype=self.dtype)
            return (2 * vs / 2000.0 ** 2).contiguous().to(self.kdev)
        vs = vd[:, g.row][:, list(g.src)].to(self.dtype)
        return ((2 * vs) / torch.full_like(vs, 2000.0 ** 2)).contiguous()

    def _scratch(self, name, shape, zero=False):
        t = self._ws.get(name)
        n = int(np.prod(shape))
        if t is None or t.numel() < n:
            self._ws.pop(name, None)
            t = torch.empty(n, dtype=self.dtype, device=self.kdev)
            self._ws[name] = t
        out = t[:n].view(*shape)
        if zero:
            out.zero_()
        return out

    def release(self):
        """Free cached device workspaces."""
        self._ws.clear()

    # ------------------------------------------------------------- kernels
    def _stream(self):
        return _lib.stream_handle(self.kdev)

    def _forward(self, geom, t0, t1, c2, amp, p0, p1, gathers=None, hist=None, ckpt=None, every=1, impl=None):
        f = _lib.fn('forward', self.dtype)
        _lib.check(f(geom, t0, t1, _lib.ptr(c2), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w),
                     _lib.ptr(amp), _lib.ptr(p0), _lib.ptr(p1), _lib.ptr(gathers), _lib.ptr(hist), _lib.ptr(ckpt),
                     every, self.impl if impl is None else impl, self._stream()), 'forward')

    def _adjoint(self, geom, t0, t1, c2, sd, res, hist, m1, m2, gacc, illum, sg, impl=None):
        f = _lib.fn('adjoint', self.dtype)
        _lib.check(f(geom, t0, t1, _lib.ptr(c2), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w),
                     _lib.ptr(sd), _lib.ptr(res), _lib.ptr(hist), _lib.ptr(m1), _lib.ptr(m2), _lib.ptr(gacc),
                     _lib.ptr(illum), _lib.ptr(sg), self.impl if impl is None else impl, self._stream()), 'adjoint')

    def _step(self, p0, p1, c2, amp, it):
        """One damped leapfrog step; returns (p2, L) (gpu.py:111-115)."""
        nb = p1.shape[0]
        a, b = p0.to(self.kdev).contiguous().clone(), p1.to(self.kdev).contiguous().clone()
        L = torch.empty(nb, self.geom.ns, 1, self.H, self.W, dtype=self.dtype, device=self.kdev)
        self._forward(self._geom(nb), it, it + 1, c2.to(self.kdev).reshape(nb, self.H, self.W).contiguous(),
                      amp.to(self.kdev).contiguous(), a, b, hist=L, impl=_lib.IMPL_STEPWISE)
        return self._out(b), self._out(L[:, :, 0])

    def forward(self, v, store=False):
        """Return gathers, optional wavefield checkpoints, and squared Courant numbers.

        Checkpoints hold the states (p0, p1) before every checkpoint_interval-th
        step, as views into one device buffer (gpu.py:117-136).  c2 has the
        reference's shape (B, 1, H, W)."""
        g = self.geom
        c2, amp = self._setup(v)
        nb = len(c2)
        shape = (nb, g.ns, self.H, self.W)
        p0 = torch.zeros(shape, dtype=self.dtype, device=self.kdev)
        p1 = torch.zeros_like(p0)
        out = torch.empty(nb, g.ns, g.nt, g.nx, dtype=self.dtype, device=self.kdev)
        ci = self.checkpoint_interval
        checkpoints = None
        if store:
            nck = -(-g.nt // ci)
            ck = torch.empty(nb, g.ns, nck, 2, self.H, self.W, dtype=self.dtype, device=self.kdev)
            self._forward(self._geom(nb), 0, g.nt, c2, amp, p0, p1, gathers=out, ckpt=ck, every=ci)
            ck = self._out(ck)
            checkpoints = {k * ci: (ck[:, :, k, 0], ck[:, :, k, 1]) for k in range(nck)}
        else:
            self._forward(self._geom(nb), 0, g.nt, c2, amp, p0, p1, gathers=out)
        return self._out(out), checkpoints, self._out(c2[:, None])

    @torch.no_grad()
    def gathers(self, v, chunk=20):
        return np.concatenate([self.forward(v[i:i + chunk])[0].cpu().numpy() for i in range(0, len(v), chunk)])

    def _check_mute(self, v_mute):
        if torch.is_tensor(v_mute):
            vm = v_mute.detach().to(self.kdev, self.dtype).reshape(-1).contiguous()
            if bool((~torch.isfinite(vm) | (vm <= 0)).any()):
                raise ValueError('mute velocities must be finite and positive')
            return vm
        v_mute = np.asarray(v_mute)
        if not np.isfinite(v_mute).all() or np.any(v_mute <= 0):
            raise ValueError('mute velocities must be finite and positive')
        return self._tensor(v_mute).reshape(-1).contiguous()

    def _mute_weights(self, vm, t_mute=0.12):
        g = self.geom
        out = torch.empty(len(vm), g.ns, g.nt, g.nx, dtype=self.dtype, device=self.kdev)
        f = _lib.fn('mute_weights', self.dtype)
        _lib.check(f(self._geom(len(vm)), _lib.ptr(vm), g.dt, g.dx, t_mute, 0.03, _lib.ptr(out), self._stream()),
                   'mute_weights')
        return out

    def mute_weights(self, v_mute, t_mute=0.12):
        """Direct-wave cosine mute, shape (len(v_mute), ns, nt, nx) (gpu.py:142-150)."""
        return self._out(self._mute_weights(self._check_mute(v_mute), t_mute))

    # --------------------------------------------------------- memory plan
    def workspace_bytes(self, nb, history=None):
        """Device bytes of one gradient/images call on nb rows under a history mode."""
        return _lib.workspace_bytes(self._geom(nb), self._segment(nb, history), self._itemsize)

    def _segment(self, nb, history=None):
        """Steps per adjoint segment: nt (whole history stored) or the checkpoint interval.
        'auto' keeps the whole history when its workspace fits 60 % of free HBM."""
        g = self.geom
        mode = self.history if history is None else history
        if mode == 'checkpoint':
            return min(self.checkpoint_interval, g.nt)
        if mode == 'full':
            return g.nt
        need = _lib.workspace_bytes(self._geom(nb), g.nt, self._itemsize)
        budget = self.max_history_bytes
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self.kdev)
            held = sum(t.numel() * t.element_size() for t in self._ws.values())
            budget = 0.6 * (free + held)
        return g.nt if need <= budget else min(self.checkpoint_interval, g.nt)

    def _driver_impl(self, seg):
        """Implementation selector of the native call: the Sim's time-loop impl for the
        whole history; for checkpoint segments the tiled kernels when asked for (or for
        small batches under 'auto'), never with 'resident'."""
        if seg >= self.geom.nt:
            return self.impl
        if self.segment_impl == 'tiled':
            return _lib.IMPL_TILED
        if self.impl != _lib.IMPL_AUTO:
            return self.impl
        return _lib.IMPL_UNTILED if self.segment_impl == 'resident' else _lib.IMPL_AUTO

    def _workspace(self, nb, seg, fresh=False):
        """Workspace of the native call: cached per (rows, segment), or a new one for a tape
        that must outlive later calls (autograd)."""
        nbytes = _lib.workspace_bytes(self._geom(nb), seg, self._itemsize)
        if fresh:
            return torch.empty(nbytes, dtype=torch.uint8, device=self.kdev)
        key = ('rtm', nb, seg)
        ws = self._ws.get(key)
        if ws is None:
            for k in [k for k in self._ws if isinstance(k, tuple) and k[0] == 'rtm']:
                del self._ws[k]
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.kdev)
            self._ws[key] = ws
        return ws

    # ----------------------------------------------------------- gradients
    def _prepare(self, v_bg):
        vd, host = self._velocity(v_bg)
        c2, amp = self._coefficients(vd, host)
        return {'c2': c2, 'amp': amp, 'sd': self._source_derivative(vd, host), 'v': vd.to(self.dtype).contiguous(),
                'nb': len(c2)}

    def _observed(self, obs, nb):
        g = self.geom
        observed = obs.detach().to(self.kdev, self.dtype) if torch.is_tensor(obs) else self._tensor(np.asarray(obs))
        if tuple(observed.shape) != (nb, g.ns, g.nt, g.nx) or not torch.isfinite(observed).all():
            raise ValueError(f'observations must be finite with shape {(nb, g.ns, g.nt, g.nx)}')
        return observed.contiguous()

    def _rtm(self, prep, stages, seg, ws, observed=None, obs_repeat=1, v_mute=None, t_mute=0.12, seed=None,
             gathers=None, grad=None, illum=None, img=None, eps=1e-4, laplacian=True):
        """One call of the native driver ss_rtm_* (forward with tape, residual, adjoint,
        fold, image) on workspace ws."""
        g = self.geom
        a = _lib.SSRtmArgs()
        a.stages, a.segment, a.impl = stages, seg, self._driver_impl(seg)
        a.obs_repeat, a.laplacian, a.timing = obs_repeat, int(bool(laplacian)), int(bool(self.timing))
        a.dt, a.dx, a.t_mute, a.ramp, a.eps, a.k = g.dt, g.dx, t_mute, 0.03, eps, (g.dt / g.dx) ** 2
        a.c2, a.A, a.B, a.w = _lib.ptr(prep['c2']), _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.w)
        a.amp, a.sd, a.v = _lib.ptr(prep['amp']), _lib.ptr(prep['sd']), _lib.ptr(prep['v'])
        a.obs, a.v_mute, a.seed = _lib.ptr(observed), _lib.ptr(v_mute), _lib.ptr(seed)
        a.gathers, a.grad, a.illum, a.img = _lib.ptr(gathers), _lib.ptr(grad), _lib.ptr(illum), _lib.ptr(img)
        fn = _lib.load().ss_rtm_f32 if self.dtype == torch.float32 else _lib.load().ss_rtm_f64
        _lib.check(fn(self._geom(prep['nb']), a, ws.data_ptr(), ws.numel(), self._stream()), 'rtm')
        if self.timing:
            self.timing_log.append(_lib.rtm_timing())

    def _forward_tape(self, prep, segment=None, fresh=False):
        """Forward pass that keeps what the adjoint needs.

        Returns (gathers, tape): the tape is the workspace holding either the whole
        Laplacian history (one adjoint launch) or checkpoints every ``segment``
        steps (the adjoint recomputes each segment, gpu.py:169-176).  ``fresh``
        puts the gathers and the tape in new tensors instead of the cached
        workspaces (for callers that keep them, e.g. autograd)."""
        g = self.geom
        nb = prep['nb']
        seg = self._segment(nb) if segment is None else min(segment, g.nt)
        ws = self._workspace(nb, seg, fresh)
        shape = (nb, g.ns, g.nt, g.nx)
        sim = torch.empty(shape, dtype=self.dtype, device=self.kdev) if fresh else self._scratch('sim', shape)
        self._rtm(prep, _lib.RTM_FORWARD, seg, ws, gathers=sim)
        return sim, {'seg': seg, 'ws': ws, 'nb': nb}

    def _adjoint_tape(self, prep, tape, res):
        """Exact adjoint of the taped forward for the data-space seed ``res``
        (nb, ns, nt, nx), imaging condition and fold (gpu.py:161-188).
        Returns (grad (nb, ns, nz, nx), illum (nb, ns, nz, nx))."""
        g = self.geom
        nb = prep['nb']
        grad = torch.empty(nb, g.ns, g.nz, g.nx, dtype=self.dtype, device=self.kdev)
        illum = torch.empty_like(grad)
        self._rtm(prep, _lib.RTM_ADJOINT, tape['seg'], tape['ws'], seed=res.contiguous(), grad=grad, illum=illum)
        return grad, illum

    def _gradient_device(self, prep, observed=None, obs_repeat=1, v_mute=None, t_mute=0.12, res=None):
        """Forward, muted residual, exact adjoint and fold on the device, one native call.

        Either ``observed`` (device, (nb/obs_repeat, ns, nt, nx)) or a ready
        data-space seed ``res`` (nb, ns, nt, nx) is given.  Returns
        (grad (nb,ns,nz,nx), illum (nb,ns,nz,nx), sim gathers)."""
        g = self.geom
        nb = prep['nb']
        seg = self._segment(nb)
        grad = torch.empty(nb, g.ns, g.nz, g.nx, dtype=self.dtype, device=self.kdev)
        illum = torch.empty_like(grad)
        sim = self._scratch('sim', (nb, g.ns, g.nt, g.nx))
        self._rtm(prep, _lib.RTM_FORWARD | _lib.RTM_ADJOINT, seg, self._workspace(nb, seg), observed=observed,
                  obs_repeat=obs_repeat, v_mute=v_mute, t_mute=t_mute,
                  seed=None if res is None else res.contiguous(), gathers=sim, grad=grad, illum=illum)
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
        prep = self._prepare(v_bg)
        observed = self._observed(obs, prep['nb'])
        vm = None if v_mute is None else self._check_mute(v_mute)
        img, illum = self.images_device(prep, observed, 1, vm, t_mute, eps, laplacian)
        return img.cpu().numpy(), illum.cpu().numpy()

    @torch.no_grad()
    def images_device(self, prep, observed, obs_repeat=1, v_mute=None, t_mute=0.12, eps=1e-4, laplacian=True):
        """Device-resident images in one native call: ``prep`` from ``_prepare``,
        ``observed`` a device tensor (nb/obs_repeat, ns, nt, nx) reused by obs_repeat
        consecutive rows, ``v_mute`` a device tensor (nb,).  Returns (img, illum).

        ``observed`` may instead be a pinned host tensor: it is then uploaded on a
        copy stream while the forward stage runs (the forward does not read the
        data), and the adjoint stage waits for the upload; the results are the same."""
        if eps <= 0:
            raise ValueError('illumination regularization eps must be positive')
        g = self.geom
        nb = prep['nb']
        seg = self._segment(nb)
        grad = self._scratch('grad', (nb, g.ns, g.nz, g.nx))
        illum = torch.empty(nb, g.ns, g.nz, g.nx, dtype=self.dtype, device=self.kdev)
        img = torch.empty_like(illum)
        ws = self._workspace(nb, seg)
        args = dict(obs_repeat=obs_repeat, v_mute=v_mute, t_mute=t_mute, grad=grad, illum=illum, img=img, eps=eps,
                    laplacian=laplacian)
        if observed.is_cuda:
            self._rtm(prep, _lib.RTM_FORWARD | _lib.RTM_ADJOINT, seg, ws, observed=observed, **args)
            return img, illum
        if not observed.is_pinned() or observed.dtype != self.dtype:
            raise ValueError('host observations must be a pinned tensor of the working dtype')
        dev = self._scratch('observed', tuple(observed.shape))
        cur = torch.cuda.current_stream(self.kdev)
        if getattr(self, '_copy_stream', None) is None:
            self._copy_stream = torch.cuda.Stream(self.kdev)
        cs = self._copy_stream
        cs.wait_stream(cur)  # the buffer's previous readers on this stream are done
        with torch.cuda.stream(cs):
            dev.copy_(observed, non_blocking=True)
            ready = torch.cuda.Event()
            ready.record(cs)
        self._rtm(prep, _lib.RTM_FORWARD, seg, ws)
        cur.wait_event(ready)
        dev.record_stream(cur)
        self._rtm(prep, _lib.RTM_ADJOINT, seg, ws, observed=dev, **args)
        return img, illum

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
        p0, p1, q0, q1 = (torch.zeros(shape, dtype=self.dtype, device=self.kdev) for _ in range(4))
        out = torch.empty(nb, g.ns, g.nt, g.nx, dtype=self.dtype, device=self.kdev)
        f = _lib.fn('born', self.dtype)
        _lib.check(f(self._geom(nb), 0, g.nt, _lib.ptr(c2), _lib.ptr(dc2), _lib.ptr(self.A), _lib.ptr(self.B),
                     _lib.ptr(self.w), _lib.ptr(amp), _lib.ptr(damp), _lib.ptr(p0), _lib.ptr(p1), _lib.ptr(q0),
                     _lib.ptr(q1), _lib.ptr(out), self._stream()), 'born')
        return self._out(out)


def calibrate(models, gathers, eps=1e-3):
    """One effective source wavelet pooled over training models with known velocity (gpu.py:213-221)."""
    s = Sim(RICKER)
    sims = s.gathers(np.asarray(models, np.float64))
    Sf = np.fft.rfft(sims, axis=2)
    Of = np.fft.rfft(np.asarray(gathers, np.float64), axis=2)
    num = (np.conj(Sf) * Of).sum(axis=(0, 1, 3))
    den = (np.abs(Sf) ** 2).sum(axis=(0, 1, 3))
    return np.fft.irfft(np.fft.rfft(RICKER) * num / (den + eps * den.max()), n=NT)
