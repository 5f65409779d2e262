"""Learned preconditioner network on hand-written HIP kernels.

A 4-layer CNN of 3x3 "same" convolutions, 2 -> 32 -> 32 -> 32 -> 1 channels,
ReLU after the first three layers (19,393 parameters).  Input: the misfit
gradient and the illumination of a model (summed over shots, each normalised
by its maximum); output: a velocity update in km/s.  Forward, backward,
the MSE loss and Adam all run in libshotsweep.so kernels; PyTorch only holds
the device buffers.  No reference counterpart: the reference describes
test-time refinement but never implemented it (README.md:25, METHOD.md:55).
"""
import numpy as np
import torch

from . import _lib

LAYERS = ((2, 32, True), (32, 32, True), (32, 32, True), (32, 1, False))


def n_params(layers=LAYERS):
    return sum(ci * co * 9 + co for ci, co, _ in layers)


class PrecondCNN:
    def __init__(self, device='cuda', seed=0, layers=LAYERS):
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise RuntimeError('PrecondCNN runs on a ROCm GPU')
        self.layers = layers
        self.n = n_params(layers)
        g = torch.Generator().manual_seed(seed)
        flat = torch.zeros(self.n, dtype=torch.float32)
        off = 0
        self.slices = []
        for ci, co, relu in layers:
            nw = co * ci * 9
            std = (2.0 / (ci * 9)) ** 0.5 if relu else (1.0 / (ci * 9)) ** 0.5 * 0.1
            flat[off:off + nw] = torch.randn(nw, generator=g) * std
            self.slices.append((off, off + nw, off + nw + co))
            off += nw + co
        self.params = flat.to(self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0

    # ------------------------------------------------------------------ views
    def _wb(self, k, buf=None):
        buf = self.params if buf is None else buf
        a, b, c = self.slices[k]
        return buf[a:b], buf[b:c]

    def _stream(self):
        return _lib.stream_handle(self.device)

    # --------------------------------------------------------------- forward
    def forward(self, x, keep=False):
        """x (N, 2, H, W) float32 device tensor -> y (N, 1, H, W); with keep, also the activations."""
        if x.dim() != 4 or x.shape[1] != self.layers[0][0]:
            raise ValueError('input must be (N, 2, H, W)')
        x = x.contiguous()
        N, _, H, W = x.shape
        lib = _lib.load()
        acts = [x]
        for k, (ci, co, relu) in enumerate(self.layers):
            w, b = self._wb(k)
            y = torch.empty(N, co, H, W, dtype=torch.float32, device=self.device)
            _lib.check(lib.ss_conv3x3_f32(N, ci, co, H, W, _lib.ptr(acts[-1]), _lib.ptr(w), _lib.ptr(b), _lib.ptr(y),
                                          int(relu), self._stream()), 'conv3x3')
            acts.append(y)
        return (acts[-1], acts) if keep else acts[-1]

    __call__ = forward

    # -------------------------------------------------------------- backward
    def backward(self, acts, dy):
        """Parameter gradients into self.grads for output gradient dy (N, 1, H, W)."""
        lib = _lib.load()
        N, _, H, W = acts[0].shape
        dz = dy.contiguous()
        for k in range(len(self.layers) - 1, -1, -1):
            ci, co, relu = self.layers[k]
            if relu:
                masked = torch.empty_like(dz)
                _lib.check(lib.ss_relu_grad_f32(dz.numel(), _lib.ptr(dz), _lib.ptr(acts[k + 1]), _lib.ptr(masked),
                                                self._stream()), 'relu_grad')
                dz = masked
            w, _ = self._wb(k)
            dw, db = self._wb(k, self.grads)
            _lib.check(lib.ss_conv3x3_wgrad_f32(N, ci, co, H, W, _lib.ptr(acts[k]), _lib.ptr(dz), _lib.ptr(dw),
                                                _lib.ptr(db), self._stream()), 'conv3x3_wgrad')
            if k > 0:
                dx = torch.empty(N, ci, H, W, dtype=torch.float32, device=self.device)
                _lib.check(lib.ss_conv3x3_dgrad_f32(N, ci, co, H, W, _lib.ptr(dz), _lib.ptr(w), _lib.ptr(dx),
                                                    self._stream()), 'conv3x3_dgrad')
                dz = dx
        return self.grads

    def loss_grad(self, pred, target):
        """0.5 * mean((pred - target)^2) and its gradient (hand-written kernel)."""
        lib = _lib.load()
        n = pred.numel()
        nblocks = int(min(1024, (n + 255) // 256))
        grad = torch.empty_like(pred)
        partial = torch.empty(nblocks, dtype=torch.float32, device=self.device)
        _lib.check(lib.ss_mse_f32(n, _lib.ptr(pred.contiguous()), _lib.ptr(target.contiguous()), _lib.ptr(grad),
                                  _lib.ptr(partial), nblocks, self._stream()), 'mse')
        return 0.5 * float(partial.double().sum()) / n, grad

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        lib = _lib.load()
        _lib.check(lib.ss_adam_f32(self.n, _lib.ptr(self.params), _lib.ptr(self.grads), _lib.ptr(self.m),
                                   _lib.ptr(self.v), lr, beta1, beta2, eps, self.step_count, self._stream()), 'adam')

    def train_step(self, x, target, lr=1e-3):
        pred, acts = self.forward(x, keep=True)
        loss, dy = self.loss_grad(pred, target)
        self.backward(acts, dy)
        self.adam_step(lr)
        return loss

    # ------------------------------------------------------------ state I/O
    def state_dict(self):
        return {'params': self.params.cpu().numpy(), 'layers': np.array(self.layers)}

    def load_state_dict(self, state):
        p = np.asarray(state['params'], np.float32)
        if p.shape != (self.n,):
            raise ValueError('parameter count mismatch')
        self.params.copy_(torch.from_numpy(p))
