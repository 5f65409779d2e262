"""ss_rtm_noalloc_f32 (the driver with its fixed device scratch instead of a
workspace) against ss_rtm_f32 with a workspace: the same kernels, so images,
illumination, gradients and gathers are bit-identical.  Two calls in a row,
the second smaller (the scratch keeps the first call's contents), plus the
size limit."""
import ctypes

import numpy as np
import pytest
import torch

from shotsweep_amd import _lib, gpu, synth
from shotsweep_amd.constants import Acquisition

pytestmark = pytest.mark.gpu


def _args(sim, prep, observed, obs_repeat, vm, seg, outs):
    g = sim.geom
    a = _lib.SSRtmArgs()
    a.stages, a.segment, a.impl = _lib.RTM_FORWARD | _lib.RTM_ADJOINT, seg, _lib.IMPL_AUTO
    a.obs_repeat, a.laplacian, a.timing = obs_repeat, 1, 0
    a.dt, a.dx, a.t_mute, a.ramp, a.eps, a.k = g.dt, g.dx, 0.12, 0.03, 1e-4, (g.dt / g.dx) ** 2
    a.c2, a.A, a.B, a.w = _lib.ptr(prep['c2']), _lib.ptr(sim.A), _lib.ptr(sim.B), _lib.ptr(sim.w)
    a.amp, a.sd, a.v = _lib.ptr(prep['amp']), _lib.ptr(prep['sd']), _lib.ptr(prep['v'])
    a.obs, a.v_mute, a.seed = _lib.ptr(observed), _lib.ptr(vm), None
    a.gathers, a.grad, a.illum, a.img = (_lib.ptr(outs[k]) for k in ('gathers', 'grad', 'illum', 'img'))
    return a


def _outs(sim, nb):
    g = sim.geom
    z = lambda *s: torch.full(s, float('nan'), dtype=torch.float32, device='cuda')
    return {'gathers': z(nb, g.ns, g.nt, g.nx), 'grad': z(nb, g.ns, g.nz, g.nx), 'illum': z(nb, g.ns, g.nz, g.nx),
            'img': z(nb, g.ns, g.nz, g.nx)}


def _case(sim, v, seg):
    """Both entry points on the same inputs; returns their outputs as NumPy arrays."""
    lib = _lib.load()
    prep = sim._prepare(v)
    nb = prep['nb']
    observed = sim.forward(v)[0].contiguous() * 0.9
    vm = torch.full((nb,), 1800.0, dtype=torch.float32, device='cuda')
    geom = sim._geom(nb)
    ref = _outs(sim, nb)
    ws = torch.empty(_lib.workspace_bytes(geom, seg, 4), dtype=torch.uint8, device='cuda')
    _lib.check(lib.ss_rtm_f32(geom, _args(sim, prep, observed, 1, vm, seg, ref), ws.data_ptr(), ws.numel(),
                              sim._stream()), 'rtm')
    got = _outs(sim, nb)
    _lib.check(lib.ss_rtm_noalloc_f32(geom, _args(sim, prep, observed, 1, vm, seg, got), sim._stream()),
               'rtm_noalloc')
    torch.cuda.synchronize()
    return {k: t.cpu().numpy() for k, t in ref.items()}, {k: t.cpu().numpy() for k, t in got.items()}


def test_noalloc_matches_workspace_entry_point_twice():
    print('\n[case] noalloc 3 rows 70x70 checkpoint 50, then 2 rows 24x28 nt 260 whole history', flush=True)
    sim = gpu.Sim(device='cuda')
    v1 = synth.mixed(1, seed=3)
    v1 = np.concatenate([v1, v1 * 1.05, v1 * 0.95]).clip(1500, 4500)
    assert _lib.workspace_bytes(sim._geom(3), 50, 4) <= 256 << 20
    ref, got = _case(sim, v1, 50)
    for k in ref:
        assert np.isfinite(ref[k]).all(), k
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    small = gpu.Sim(device='cuda', pad=12, geom=Acquisition(nz=24, nx=28, nt=260, ns=3))
    v2 = 1800 + 900 * np.random.default_rng(4).random((2, 24, 28))
    ref, got = _case(small, v2, 260)
    for k in ref:
        assert np.isfinite(ref[k]).all(), k
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_noalloc_rejects_calls_beyond_its_scratch():
    print('\n[case] noalloc size limit', flush=True)
    sim = gpu.Sim(device='cuda')
    nb = 192
    geom = sim._geom(nb)
    assert _lib.workspace_bytes(geom, 1000, 4) > 256 << 20
    a = _lib.SSRtmArgs()
    a.stages, a.segment = _lib.RTM_FORWARD | _lib.RTM_ADJOINT, 1000
    rc = _lib.load().ss_rtm_noalloc_f32(geom, a, sim._stream())
    assert rc == -1 and b'fixed scratch' in _lib.load().ss_last_error()
