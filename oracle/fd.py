"""Eager-PyTorch CPU restatement of the acoustic FD simulator, its exact adjoint and RTM images.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker for the HIP
path and the timed host-CPU baseline, never part of the product.

Follows ``openfwi/physics/gpu.py`` (all of lines 13-221) and
``openfwi/physics/constants.py:4-7`` operation for operation, so that at the
reference geometry (70 x 70 cells, pad 40, 5 shots, 1000 steps) its FP32 and
FP64 results are those of the reference's own CPU path:

* every elementwise expression is evaluated in the same order as the reference
  (the rounding of each FP32 operation is part of the contract);
* coefficients are formed in float64 with NumPy and rounded once to the
  working dtype, as ``Sim.__init__``/``Sim._setup`` do;
* the checkpoint-and-recompute adjoint keeps the per-segment Laplacian list.

Generalisations (no reference counterpart, used for uneven batches):
rectangular grids (separate damping profiles in z and x), any number of
evenly spaced shots, any number of steps, receivers on every column.

Parity pins for each function (tests/test_oracle.py):
  lap, lap_adjoint        closed form on quadratics, transpose identity, brute force
  fold_edge_padding       transpose identity against np.pad(mode='edge')
  damping/ricker/setup    closed forms
  forward                 reference golden fixtures (FP32/FP64), linearity in the
                          wavelet, mirror symmetry, reciprocity
  _gradient / images      reference golden fixtures, finite differences,
                          checkpoint-interval invariance, dot-product test with born()
  calibrate               reference golden fixture, Wiener closed form
  born                    finite differences of forward (no reference counterpart)
"""
import math
from dataclasses import dataclass, field

import numpy as np
import torch

# openfwi/physics/constants.py:4-7
DX = 10.0
DT = 1e-3
F0 = 15.0
VMIN, VMAX = 1500.0, 4500.0
DAMP_VMAX = 4500.0            # gpu.py:88 hard-codes 4500 m/s in the sponge strength


def even_sources(nx, ns):
    """Evenly spaced source columns, rounded half-to-even.

    At nx=70, ns=5 this is [0, 17, 34, 52, 69], the reference's SRC
    (constants.py:6 SRC_X / DX, gpu.py:14)."""
    return [int(c) for c in np.round(np.linspace(0, nx - 1, ns))]


@dataclass(frozen=True)
class Geometry:
    """Acquisition of one propagation family (constants.py:4-7 generalised)."""
    nz: int = 70
    nx: int = 70
    nt: int = 1000
    ns: int = 5
    dx: float = DX
    dt: float = DT
    f0: float = F0
    row: int = 1                       # source and receiver model row (10 m datum)
    src: tuple = field(default=None)   # source columns (model coordinates)

    def __post_init__(self):
        if self.src is None:
            object.__setattr__(self, 'src', tuple(even_sources(self.nx, self.ns)))
        if len(self.src) != self.ns:
            raise ValueError('one source column per shot is required')

    @property
    def src_x(self):
        """Source positions in metres (constants.py:6)."""
        return np.asarray(self.src) * self.dx


REFERENCE = Geometry()


def ricker(geom=REFERENCE):
    """gpu.py:15-17."""
    t = np.arange(geom.nt) * geom.dt
    a = (np.pi * geom.f0 * (t - 1.2 / geom.f0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


RICKER = ricker()


def damping(geom, pad, R):
    """Quadratic sponge coefficients A = 1/(1+a), B = 1-a in float64 (gpu.py:85-90)."""
    def profile(n):
        i = np.arange(n)
        return np.clip(np.maximum(pad - i, i - (n - 1 - pad)), 0, None) / pad
    dz = profile(geom.nz + 2 * pad)
    dxp = profile(geom.nx + 2 * pad)
    a = 3 * DAMP_VMAX / (2 * pad * geom.dx) * np.log(1 / R) * (dz[:, None] ** 2 + dxp[None, :] ** 2) * geom.dt / 2
    return 1 / (1 + a), 1 - a


def lap(p):
    """Fourth-order Laplacian times dx^2, zero on the two outer cells (gpu.py:33-37).

    The sum is evaluated left to right in the reference's order: the z taps
    (-2, -1, 0, +1, +2), then the x taps, each centre weight -30 entering once
    per axis, then one division by 12."""
    out = torch.zeros_like(p)
    mid = p[..., 2:-2, 2:-2]
    zm2, zm1, zp1, zp2 = p[..., :-4, 2:-2], p[..., 1:-3, 2:-2], p[..., 3:-1, 2:-2], p[..., 4:, 2:-2]
    xm2, xm1, xp1, xp2 = p[..., 2:-2, :-4], p[..., 2:-2, 1:-3], p[..., 2:-2, 3:-1], p[..., 2:-2, 4:]
    acc = -zm2 + 16 * zm1 - 30 * mid + 16 * zp1 - zp2 - xm2 + 16 * xm1 - 30 * mid + 16 * xp1 - xp2
    out[..., 2:-2, 2:-2] = acc / 12
    return out


# Scatter taps of the transposed stencil, in the reference's accumulation order
# (gpu.py:44-52): (row shift of the destination, column shift, weight).
_ADJ_TAPS = ((0, 0, -60), (-2, 0, -1), (-1, 0, 16), (1, 0, 16), (2, 0, -1),
             (0, -2, -1), (0, -1, 16), (0, 1, 16), (0, 2, -1))


def lap_adjoint(p):
    """Exact transpose of ``lap`` including its zero two-cell border (gpu.py:40-53)."""
    out = torch.zeros_like(p)
    q = p[..., 2:-2, 2:-2] / 12
    nz, nx = p.shape[-2:]
    for dz, dx, weight in _ADJ_TAPS:
        dst = out[..., 2 + dz:nz - 2 + dz, 2 + dx:nx - 2 + dx]
        if weight == -1:
            dst -= q
        elif weight < 0:
            dst -= (-weight) * q
        else:
            dst += weight * q
    return out


def fold_edge_padding(g, pad, nz=70, nx=70):
    """Transpose of np.pad(mode='edge') over the last two axes (gpu.py:56-67)."""
    p = pad
    out = g[..., p:p + nz, p:p + nx].clone()
    out[..., 0, :] += g[..., :p, p:p + nx].sum(-2)
    out[..., -1, :] += g[..., p + nz:, p:p + nx].sum(-2)
    out[..., :, 0] += g[..., p:p + nz, :p].sum(-1)
    out[..., :, -1] += g[..., p:p + nz, p + nx:].sum(-1)
    for zi, zs in ((0, slice(None, p)), (-1, slice(p + nz, None))):
        for xi, xs in ((0, slice(None, p)), (-1, slice(p + nx, None))):
            out[..., zi, xi] += g[..., zs, xs].sum((-2, -1))
    return out


def imaging_memory_bytes(batch, pad=40, checkpoint_interval=40, bytes_per_value=4, geom=REFERENCE):
    """Tensor-storage estimate of one imaging call (gpu.py:70-73)."""
    fields = 2 * math.ceil(geom.nt / checkpoint_interval) + checkpoint_interval + 24
    return int(batch * geom.ns * bytes_per_value * ((geom.nz + 2 * pad) * (geom.nx + 2 * pad) * fields
                                                     + 3 * geom.nt * geom.nx))


class OracleSim:
    """CPU restatement of ``gpu.Sim`` (gpu.py:76-210) on a generalised geometry."""

    def __init__(self, wavelet=None, pad=40, R=1e-5, checkpoint_interval=40,
                 dtype=torch.float32, geom=REFERENCE):
        if pad < 2 or checkpoint_interval < 1:
            raise ValueError('pad must be >= 2 and checkpoint_interval must be positive')
        self.geom = geom
        self.pad = pad
        self.checkpoint_interval = checkpoint_interval
        self.dtype = dtype
        A, B = damping(geom, pad, R)
        self.A = self._tensor(A)
        self.B = self._tensor(B)
        self.wav = np.asarray(ricker(geom) if wavelet is None else wavelet)
        if self.wav.shape != (geom.nt,) or not np.isfinite(self.wav).all() or not np.any(self.wav):
            raise ValueError(f'wavelet must contain {geom.nt} finite samples and nonzero energy')
        self.w = self._tensor(self.wav * 1e3)
        self.si = torch.arange(geom.ns)
        self.sx = torch.tensor(geom.src) + pad

    def _tensor(self, value):
        return torch.as_tensor(value, dtype=self.dtype)

    @property
    def shape(self):
        return self.geom.nz + 2 * self.pad, self.geom.nx + 2 * self.pad

    def _setup(self, v):
        """c2 = vp^2 (dt/dx)^2 and source amplitude (v/2000)^2 (gpu.py:101-109)."""
        g = self.geom
        v = np.asarray(v, np.float64)
        if v.ndim != 3 or v.shape[1:] != (g.nz, g.nx) or not np.isfinite(v).all() or np.any(v <= 0):
            raise ValueError(f'velocities must be finite positive (B, {g.nz}, {g.nx}) arrays')
        p = self.pad
        vp = np.pad(v, ((0, 0), (p, p), (p, p)), mode='edge')
        c2 = self._tensor(vp ** 2 * (g.dt / g.dx) ** 2)[:, None]
        amp = self._tensor((v[:, g.row, list(g.src)] / 2000.0) ** 2)
        return c2, amp

    def _step(self, p0, p1, c2, amp, it):
        """Damped leapfrog step, source added after damping (gpu.py:111-115)."""
        L = lap(p1)
        p2 = (2 * p1 - self.B * p0 + c2 * L) * self.A
        p2[:, self.si, self.pad + self.geom.row, self.sx] += self.w[it] * amp
        return p2, L

    def forward(self, v, store=False):
        """Gathers (B, ns, nt, nx), optional checkpoints, c2 (gpu.py:117-136)."""
        g, p = self.geom, self.pad
        c2, amp = self._setup(v)
        H, W = self.shape
        p0 = torch.zeros(len(v), g.ns, H, W, dtype=self.dtype)
        p1 = torch.zeros_like(p0)
        out = torch.zeros(len(v), g.ns, g.nt, g.nx, dtype=self.dtype)
        checkpoints = {} if store else None
        for it in range(g.nt):
            if store and it % self.checkpoint_interval == 0:
                checkpoints[it] = (p0.clone(), p1.clone())
            p2, _ = self._step(p0, p1, c2, amp, it)
            out[:, :, it] = p2[:, :, p + g.row, p:p + g.nx]
            p0, p1 = p1, p2
        return out, checkpoints, c2

    @torch.no_grad()
    def gathers(self, v, chunk=20):
        """gpu.py:138-140."""
        return np.concatenate([self.forward(v[i:i + chunk])[0].numpy() for i in range(0, len(v), chunk)])

    def mute_weights(self, v_mute, t_mute=0.12):
        """Direct-wave cosine mute with a 30 ms ramp (gpu.py:142-150)."""
        g = self.geom
        v_mute = np.asarray(v_mute)
        if not np.isfinite(v_mute).all() or np.any(v_mute <= 0):
            raise ValueError('mute velocities must be finite and positive')
        h = self._tensor(np.abs(np.arange(g.nx)[None, :] * g.dx - np.asarray(g.src_x)[:, None]))
        t = torch.arange(g.nt, dtype=self.dtype)[:, None] * g.dt
        t0 = h[None] / self._tensor(v_mute)[:, None, None] + t_mute
        ramp = torch.clamp((t[None, None] - t0[:, :, None, :]) / 0.03, 0, 1)
        return 0.5 - 0.5 * torch.cos(np.pi * ramp)

    def residual(self, sim, obs, v_mute=None, t_mute=0.12):
        """res = (sim - obs) * mute (gpu.py:155-160)."""
        observed = self._tensor(np.asarray(obs))
        if observed.shape != sim.shape or not torch.isfinite(observed).all():
            raise ValueError(f'observations must be finite with shape {tuple(sim.shape)}')
        res = sim - observed
        if v_mute is not None:
            res *= self.mute_weights(v_mute, t_mute)
        return res

    def adjoint(self, v_bg, res, checkpoints, c2):
        """Checkpointed exact adjoint + imaging condition + epilogue (gpu.py:161-188).

        ``res`` is the data-space seed (B, ns, nt, nx); returns (grad, illum)."""
        g, p = self.geom, self.pad
        _, amp = self._setup(v_bg)
        H, W = self.shape
        shape = (len(v_bg), g.ns, H, W)
        m1 = torch.zeros(shape, dtype=self.dtype)
        m2 = torch.zeros_like(m1)
        grad_acc = torch.zeros_like(m1)
        source_grad = torch.zeros(len(v_bg), g.ns, dtype=self.dtype)
        illum = torch.zeros(len(v_bg), g.ns, g.nz, g.nx, dtype=self.dtype)
        source_derivative = 2 * self._tensor(np.asarray(v_bg)[:, g.row, list(g.src)]) / 2000.0 ** 2
        r = p + g.row
        for start in reversed(list(checkpoints)):
            p0, p1 = checkpoints.pop(start)
            end = min(start + self.checkpoint_interval, g.nt)
            laplacians = []
            for it in range(start, end):
                p2, L = self._step(p0, p1, c2, amp, it)
                laplacians.append(L)
                p0, p1 = p1, p2
            for it in range(end - 1, start - 1, -1):
                L = laplacians.pop()
                mu = 2 * self.A * m1 + lap_adjoint(c2 * self.A * m1) - self.A * self.B * m2
                mu[:, :, r, p:p + g.nx] += res[:, :, it]
                grad_acc += self.A * mu * L
                source_grad += mu[:, self.si, r, self.sx] * self.w[it] * source_derivative
                illum += L[..., p:p + g.nz, p:p + g.nx] ** 2
                m2, m1 = m1, mu
        vp = self._tensor(np.pad(np.asarray(v_bg), ((0, 0), (p, p), (p, p)), mode='edge'))[:, None]
        scaled = grad_acc * 2 * vp * (g.dt / g.dx) ** 2
        self.last_scaled = scaled      # the unfolded terms, kept for rounding bounds in tests
        grad = fold_edge_padding(scaled, p, g.nz, g.nx)
        grad[:, self.si, g.row, torch.tensor(list(g.src))] += source_grad
        return grad, illum

    def _gradient(self, v_bg, obs, v_mute=None, t_mute=0.12):
        """gpu.py:152-188."""
        sim, checkpoints, c2 = self.forward(v_bg, store=True)
        res = self.residual(sim, obs, v_mute, t_mute)
        return self.adjoint(v_bg, res, checkpoints, c2)

    @torch.no_grad()
    def gradient(self, v_bg, obs, v_mute=None, t_mute=0.12):
        """Per-shot derivative of .5*sum(w*(sim-obs)^2) (gpu.py:190-194)."""
        grad, illum = self._gradient(v_bg, obs, v_mute, t_mute)
        return grad.numpy(), illum.numpy()

    @staticmethod
    def image_epilogue(grad, illum, eps=1e-4, laplacian=True):
        """Illumination divide and Laplacian filter (gpu.py:206-209)."""
        denominator = illum + eps * illum.amax(dim=(2, 3), keepdim=True)
        img = -grad / denominator.clamp_min(torch.finfo(grad.dtype).tiny)
        if laplacian:
            img = -lap(torch.nn.functional.pad(img, (2, 2, 2, 2), mode='replicate'))[..., 2:-2, 2:-2]
        return img

    @torch.no_grad()
    def images(self, v_bg, obs, v_mute=None, t_mute=0.12, eps=1e-4, laplacian=True):
        """gpu.py:196-210."""
        if eps <= 0:
            raise ValueError('illumination regularization eps must be positive')
        grad, illum = self._gradient(v_bg, obs, v_mute, t_mute)
        return self.image_epilogue(grad, illum, eps, laplacian).numpy(), illum.numpy()

    @torch.no_grad()
    def born(self, v, dv):
        """Tangent-linear (Born) gathers: d(gathers)/dv applied to dv.

        No reference counterpart. Obtained by differentiating ``_step`` and
        ``_setup``: dc2 = 2 vp dvp (dt/dx)^2 with dvp the edge-padded dv, and
        damp(amp) = 2 v[row, src] dv[row, src] / 2000^2 at the source."""
        g, p = self.geom, self.pad
        c2, amp = self._setup(v)
        vv = np.asarray(v, np.float64)
        dvv = np.asarray(dv, np.float64)
        vp = np.pad(vv, ((0, 0), (p, p), (p, p)), mode='edge')
        dvp = np.pad(dvv, ((0, 0), (p, p), (p, p)), mode='edge')
        dc2 = self._tensor(2 * vp * dvp * (g.dt / g.dx) ** 2)[:, None]
        damp = self._tensor(2 * vv[:, g.row, list(g.src)] * dvv[:, g.row, list(g.src)] / 2000.0 ** 2)
        H, W = self.shape
        p0 = torch.zeros(len(v), g.ns, H, W, dtype=self.dtype)
        p1 = torch.zeros_like(p0)
        q0 = torch.zeros_like(p0)
        q1 = torch.zeros_like(p0)
        out = torch.zeros(len(v), g.ns, g.nt, g.nx, dtype=self.dtype)
        for it in range(g.nt):
            L = lap(p1)
            q2 = (2 * q1 - self.B * q0 + c2 * lap(q1) + dc2 * L) * self.A
            q2[:, self.si, p + g.row, self.sx] += self.w[it] * damp
            p2, _ = self._step(p0, p1, c2, amp, it)
            out[:, :, it] = q2[:, :, p + g.row, p:p + g.nx]
            p0, p1, q0, q1 = p1, p2, q1, q2
        return out.numpy()


def calibrate(models, gathers, eps=1e-3, geom=REFERENCE):
    """Wiener-filtered effective source wavelet (gpu.py:213-221)."""
    base = ricker(geom)
    s = OracleSim(base, geom=geom)
    sims = s.gathers(np.asarray(models, np.float64))
    Sf = np.fft.rfft(sims, axis=2)
    Of = np.fft.rfft(np.asarray(gathers, np.float64), axis=2)
    num = (np.conj(Sf) * Of).sum(axis=(0, 1, 3))
    den = (np.abs(Sf) ** 2).sum(axis=(0, 1, 3))
    return np.fft.irfft(np.fft.rfft(base) * num / (den + eps * den.max()), n=geom.nt)
