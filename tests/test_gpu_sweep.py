"""Mixed-size tiled sweep (ss_sweep_*_f32) against the stepwise kernels and the oracle.

The tiled kernels recompute halo cells with the same per-cell arithmetic, so
every output (gathers, Laplacian history, imaging sum, illumination, source
term, images) must equal the stepwise FP32 path bit for bit, for any tiling:
tiles that do not divide the grid, depths that do not divide nt, one tile per
plane and one step per launch.  The stepwise path itself is pinned to the
reference's golden vectors (test_gpu_parity.py); the images are also checked
against the oracle at 1e-5 relative L2.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import fd
from shotsweep_amd import _lib, gpu, synth
from shotsweep_amd.constants import Acquisition
from shotsweep_amd.mixed import ADJOINT, FORWARD, MixedSweep, config_acquisitions

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda')
PAD = 10
# long enough (0.23-0.26 s) for the residual to survive the 0.12 s direct-wave mute
ACQS = [Acquisition(nz=20, nx=30, nt=261, ns=3), Acquisition(nz=26, nx=22, nt=235, ns=2),
        Acquisition(nz=16, nx=40, nt=250, ns=4, row=3)]
# odd plane widths (nx + 2 pad) run the cell kernels; the even ones above the paired kernels
ACQS_ODD = [Acquisition(nz=20, nx=31, nt=261, ns=3), Acquisition(nz=26, nx=23, nt=235, ns=2),
            Acquisition(nz=16, nx=41, nt=250, ns=4, row=3)]
NB = (2, 3, 1)
TILINGS = [(8, 12, 1), (13, 17, 3), (60, 60, 2), (11, 64, 4), (30, 9, 5)]


def case(name):
    print(f'\n[case] {name}', flush=True)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(autouse=True)
def sync_after():
    yield
    torch.cuda.synchronize()


def models(acq, n, seed):
    return synth.mixed(n, seed=seed, nz=acq.nz, nx=acq.nx)


def stepwise_forward(sim, prep):
    g = sim.geom
    nb = prep['nb']
    p0 = torch.zeros(nb, g.ns, sim.H, sim.W, device=DEV)
    p1 = torch.zeros_like(p0)
    out = torch.empty(nb, g.ns, g.nt, g.nx, device=DEV)
    hist = torch.empty(nb, g.ns, g.nt, sim.H, sim.W, device=DEV)
    sim._forward(sim._geom(nb), 0, g.nt, prep['c2'], prep['amp'], p0, p1, gathers=out, hist=hist,
                 impl=_lib.IMPL_STEPWISE)
    return out, hist


def stepwise_adjoint(sim, prep, res, hist):
    g = sim.geom
    nb = prep['nb']
    z = lambda *s: torch.zeros(*s, device=DEV)
    m1, m2, gacc = z(nb, g.ns, sim.H, sim.W), z(nb, g.ns, sim.H, sim.W), z(nb, g.ns, sim.H, sim.W)
    illum, sg = z(nb, g.ns, g.nz, g.nx), z(nb, g.ns)
    sim._adjoint(sim._geom(nb), 0, g.nt, prep['c2'], prep['sd'], res, hist, m1, m2, gacc, illum, sg,
                 impl=_lib.IMPL_STEPWISE)
    return gacc, illum, sg


def setup_parts(ms, seed=0):
    parts = []
    for k, (sim, nb) in enumerate(zip(ms.sims, NB)):
        prep = sim._prepare(models(sim.geom, nb, seed + k))
        parts.append((k, prep, ms._buffers(k, nb)))
    return parts


@pytest.mark.parametrize('layout', ['pairs', 'cells'])
@pytest.mark.parametrize('tiling', TILINGS)
def test_sweep_forward_and_adjoint_bit_identical_to_stepwise(tiling, layout):
    case(f'sweep tiling={tiling} {layout}')
    acqs = ACQS if layout == 'pairs' else ACQS_ODD
    tiles = [{FORWARD: tiling, ADJOINT: tiling} for _ in acqs]
    ms = MixedSweep(acqs, pad=PAD, autotune=False, use_graph=False, tiles=tiles)
    parts = setup_parts(ms)
    ms._run(parts, FORWARD)
    gen = torch.Generator(device=DEV).manual_seed(3)
    refs = []
    for k, prep, bufs in parts:
        sim = ms.sims[k]
        out, hist = stepwise_forward(sim, prep)
        assert torch.equal(bufs['sim'], out), f'gathers group {k}'
        assert torch.equal(bufs['hist'], hist), f'history group {k}'
        bufs['res'].copy_(torch.randn(bufs['res'].shape, generator=gen, device=DEV))
        refs.append(stepwise_adjoint(sim, prep, bufs['res'], hist))
    ms._run(parts, ADJOINT)
    for (k, prep, bufs), (gacc, illum, sg) in zip(parts, refs):
        assert torch.equal(bufs['gacc'], gacc), f'imaging sum group {k}'
        assert torch.equal(bufs['illum'], illum), f'illumination group {k}'
        assert torch.equal(bufs['sg'], sg), f'source term group {k}'


def small_batches(seed=0, rep=2):
    batches, true = [], []
    for k, acq in enumerate(ACQS):
        sim = gpu.Sim(geom=acq, pad=PAD)
        m = models(acq, NB[k], seed + 10 * k)
        obs = sim.forward(m)[0].contiguous()
        bg = np.repeat(synth.smooth(m, 4.0), rep, 0)
        vm = torch.tensor(np.repeat(m[:, 0].mean(1), rep), dtype=torch.float32, device=DEV)
        batches.append({'v_bg': bg, 'obs': obs, 'obs_repeat': rep, 'v_mute': vm})
        true.append(m)
    return batches, true


def test_mixed_images_equal_single_size_sims_and_oracle():
    case('mixed images vs Sim and oracle')
    batches, _ = small_batches()
    ms = MixedSweep(ACQS, pad=PAD)
    out = ms.images_device(batches)
    for k, (acq, b) in enumerate(zip(ACQS, batches)):
        sim = gpu.Sim(geom=acq, pad=PAD, impl=_lib.IMPL_STEPWISE)
        prep = sim._prepare(b['v_bg'])
        img, illum = sim.images_device(prep, b['obs'], b['obs_repeat'], b['v_mute'])
        np.testing.assert_array_equal(out[k][0].cpu().numpy(), img.cpu().numpy())
        np.testing.assert_array_equal(out[k][1].cpu().numpy(), illum.cpu().numpy())
        osim = fd.OracleSim(pad=PAD, geom=fd.Geometry(nz=acq.nz, nx=acq.nx, nt=acq.nt, ns=acq.ns, row=acq.row))
        obs_rep = np.repeat(b['obs'].cpu().numpy(), b['obs_repeat'], 0)
        oimg, oill = osim.images(b['v_bg'], obs_rep, v_mute=b['v_mute'].cpu().numpy())
        assert np.abs(oimg).max() > 0
        assert rel(out[k][0].cpu().numpy(), oimg) < 1e-5
        np.testing.assert_array_equal(out[k][1].cpu().numpy(), oill)
    for k in range(len(ACQS)):
        th, tw, d = ms.tiles[k][FORWARD]
        assert th >= 1 and tw >= 1 and d >= 1


def test_graph_replays_and_chunks_match_direct_launches():
    case('graph replay / chunking')
    batches, _ = small_batches(seed=5)
    ref = MixedSweep(ACQS, pad=PAD, autotune=False, use_graph=False).images_device(batches)
    ms = MixedSweep(ACQS, pad=PAD, autotune=False, use_graph=True)
    for rep in range(3):
        out = ms.images_device(batches)
        x = torch.randn(1000, device=DEV)
        _ = float((x * x).sum())       # eager work between replays
        torch.cuda.synchronize()
        for k in range(len(ACQS)):
            assert torch.equal(out[k][0], ref[k][0]), f'replay {rep} group {k}'
            assert torch.equal(out[k][1], ref[k][1])
    small = MixedSweep(ACQS, pad=PAD, autotune=False, max_history_bytes=ms.history_bytes([4, 6, 2]) // 3)
    n, _ = small._chunks([4, 6, 2], [2, 2, 2])
    assert n >= 2
    out = small.images_device(batches)
    for k in range(len(ACQS)):
        assert torch.equal(out[k][0], ref[k][0])
        assert torch.equal(out[k][1], ref[k][1])


def test_uneven_workload_sizes_bit_identical_to_stepwise():
    """One model per grid size of the uneven-batch workload (70x70/100x140/200x200,
    nt 1000/1500/2000, 5/7/10 shots, pad 40) under the autotuned tilings."""
    case('uneven workload sizes vs stepwise')
    acqs = config_acquisitions()
    batches = []
    for k, acq in enumerate(acqs):
        sim = gpu.Sim(geom=acq)
        m = models(acq, 1, 40 + k)
        obs = sim.forward(m)[0].contiguous()
        vm = torch.tensor(m[:, 0].mean(1), dtype=torch.float32, device=DEV)
        batches.append({'v_bg': synth.smooth(m, 8.0), 'obs': obs, 'obs_repeat': 1, 'v_mute': vm})
    ms = MixedSweep(acqs)
    out = ms.images_device(batches)
    for k, (acq, b) in enumerate(zip(acqs, batches)):
        sim = gpu.Sim(geom=acq, impl=_lib.IMPL_STEPWISE)
        prep = sim._prepare(b['v_bg'])
        img, illum = sim.images_device(prep, b['obs'], 1, b['v_mute'])
        print(f'  group {k}: tiles {ms.tiles[k]}', flush=True)
        assert torch.equal(out[k][0], img), f'images group {k}'
        assert torch.equal(out[k][1], illum), f'illumination group {k}'


def _sizes_fixture():
    import importlib.util
    import sys
    from pathlib import Path
    here = Path(__file__).resolve().parent / 'golden'
    spec = importlib.util.spec_from_file_location('make_oracle_sizes', here / 'make_oracle_sizes.py')
    mod = importlib.util.module_from_spec(spec)
    sys.modules.setdefault('make_oracle_sizes', mod)
    spec.loader.exec_module(mod)
    with np.load(here / 'oracle_uneven_sizes.npz') as d:
        return mod, {k: d[k] for k in d.files}


@pytest.mark.parametrize('size', [0, 1])
def test_uneven_sizes_match_oracle_at_full_size(size):
    """One background row at each larger grid of the uneven batch at its real size
    (100x140, nt 1500, 7 shots; 200x200, nt 2000, 10 shots) against the oracle's
    fixture (tests/golden/make_oracle_sizes.py): gathers and illumination the same
    bytes (SHA-256), images within 1e-5 relative L2, through both the drop-in
    Sim and the mixed sweep that runs the uneven batch."""
    mod, fix = _sizes_fixture()
    nz, nx, nt, ns, seed = mod.SIZES[size]
    tag = f'{nz}x{nx}'
    case(f'uneven size {tag} nt {nt} ns {ns} vs oracle fixture')
    model, bg, vm = mod.case_inputs(nz, nx, seed)
    acq = Acquisition(nz=nz, nx=nx, nt=nt, ns=ns)
    sim = gpu.Sim(geom=acq)
    obs = sim.forward(model)[0].contiguous()
    assert mod.digest(obs.cpu().numpy()) == str(fix[f'{tag}_gathers_sha256'])
    img, illum = sim.images(bg, obs.cpu().numpy(), v_mute=vm)
    assert mod.digest(illum) == str(fix[f'{tag}_illum_sha256'])
    e = rel(img, fix[f'{tag}_img'])
    ms = MixedSweep([acq])
    (mimg, mill), = ms.images_device([{'v_bg': bg, 'obs': obs, 'obs_repeat': 1,
                                       'v_mute': torch.tensor(vm, dtype=torch.float32, device=DEV)}])
    assert mod.digest(mill.cpu().numpy()) == str(fix[f'{tag}_illum_sha256'])
    em = rel(mimg.cpu().numpy(), fix[f'{tag}_img'])
    print(f'  image rel L2 vs oracle: Sim {e:.3e}, mixed sweep {em:.3e}')
    assert e < 1e-5 and em < 1e-5


def _forward_group(sim, prep, nb, state, hist, gathers, tiles=(0, 0, 0)):
    g = _lib.SSSweepGroup()
    g.geom = sim._geom(nb)
    g.tile_h, g.tile_w, g.depth = tiles
    g.c2, g.A, g.B, g.w = _lib.ptr(prep['c2']), _lib.ptr(sim.A), _lib.ptr(sim.B), _lib.ptr(sim.w)
    g.amp, g.state = _lib.ptr(prep['amp']), _lib.ptr(state)
    g.hist, g.gathers = _lib.ptr(hist), _lib.ptr(gathers)
    return (_lib.SSSweepGroup * 1)(g)


@pytest.mark.parametrize('shape', [(200, 200, 10, 7, (94, 94, 8), 300), (100, 140, 7, 19, (60, 110, 6), 301)])
def test_band_forward_bit_identical_to_stepwise(shape):
    """The band kernels, which hold a plane in a team of workgroups that trade
    boundary rows every step: 280 x 280 planes (200 x 200 models, pad 40) in four
    bands, 7 rows x 10 shots = 70 propagations, two rounds of the 64 teams; 180 x
    220 planes (100 x 140 models) in two bands, 19 rows x 7 shots = 133
    propagations, two rounds of the 128 teams (the second rounds partial); nt 300
    and 301 (an odd step count: the time loop's last step outside its pairs).
    Gathers, the whole Laplacian history (zero on the rim) and the final state
    equal the stepwise kernels' bit for bit; the state buffers are NaN on entry,
    which a zero-state call must not read.  The tiled kernels under a forced
    tiling of the same grid give the same gathers and history."""
    nz, nx, ns, nb, tiling, nt = shape
    acq = Acquisition(nz=nz, nx=nx, nt=nt, ns=ns)
    case(f'band forward {acq.nz}x{acq.nx} nt {acq.nt} ns {acq.ns} x {nb} rows vs stepwise')
    sim = gpu.Sim(geom=acq)
    g = acq
    prep = sim._prepare(models(acq, nb, 11))
    H, W = sim.H, sim.W
    p0 = torch.zeros(nb, g.ns, H, W, device=DEV)
    p1 = torch.zeros_like(p0)
    ref_out = torch.empty(nb, g.ns, g.nt, g.nx, device=DEV)
    ref_hist = torch.empty(nb, g.ns, g.nt, H, W, device=DEV)
    sim._forward(sim._geom(nb), 0, g.nt, prep['c2'], prep['amp'], p0, p1, gathers=ref_out, hist=ref_hist,
                 impl=_lib.IMPL_STEPWISE)
    lib = _lib.load()
    for tiles in [(0, 0, 0), tiling]:
        state = torch.full((4, nb, g.ns, H, W), float('nan'), device=DEV)
        hist = torch.full_like(ref_hist, float('nan'))
        out = torch.full_like(ref_out, float('nan'))
        arr = _forward_group(sim, prep, nb, state, hist, out, tiles)
        _lib.check(lib.ss_sweep_forward_f32(1, arr, 1, _lib.stream_handle(DEV)), 'sweep forward')
        torch.cuda.synchronize()
        assert torch.equal(out, ref_out), f'gathers, tiles {tiles}'
        assert torch.equal(hist, ref_hist), f'history, tiles {tiles}'
        if tiles == (0, 0, 0):  # one launch from state buffer 0: the final state in buffer 1
            assert torch.equal(state[2], p0) and torch.equal(state[3], p1)
        print(f'  tiles {tiles}: equal', flush=True)


def _adjoint_group(sim, prep, nb, state, res, hist, gacc, illum, sg, tiles=(0, 0, 0)):
    g = _lib.SSSweepGroup()
    g.geom = sim._geom(nb)
    g.tile_h, g.tile_w, g.depth = tiles
    g.c2, g.A, g.B, g.w = _lib.ptr(prep['c2']), _lib.ptr(sim.A), _lib.ptr(sim.B), _lib.ptr(sim.w)
    g.amp, g.res, g.state, g.hist = _lib.ptr(prep['sd']), _lib.ptr(res), _lib.ptr(state), _lib.ptr(hist)
    g.gacc, g.illum, g.src_grad = _lib.ptr(gacc), _lib.ptr(illum), _lib.ptr(sg)
    return (_lib.SSSweepGroup * 1)(g)


@pytest.mark.parametrize('shape', [(200, 200, 10, 7, (70, 94, 6), 300), (100, 140, 7, 19, (90, 74, 6), 301)])
def test_band_adjoint_bit_identical_to_stepwise(shape):
    """The adjoint on the band kernels (the bands trade the new state of their edge
    rows every step and form the halo's q from it) at the same two shapes and
    propagation counts as the forward test, nt 300 and 301 (an odd step count: the
    time loop's last step outside its pairs), with
    the stepwise forward's Laplacian history and its gathers as the residual:
    imaging sum, illumination, source term and the final state (m1, m2) equal the
    stepwise adjoint's bit for bit; the state buffers and outputs are NaN on
    entry.  The tiled kernels under a forced tiling give the same imaging sum,
    illumination and source term."""
    nz, nx, ns, nb, tiling, nt = shape
    acq = Acquisition(nz=nz, nx=nx, nt=nt, ns=ns)
    case(f'band adjoint {acq.nz}x{acq.nx} nt {acq.nt} ns {acq.ns} x {nb} rows vs stepwise')
    sim = gpu.Sim(geom=acq)
    g = acq
    prep = sim._prepare(models(acq, nb, 12))
    H, W = sim.H, sim.W
    res, hist = stepwise_forward(sim, prep)
    z = lambda *s: torch.zeros(*s, device=DEV)
    m1, m2 = z(nb, g.ns, H, W), z(nb, g.ns, H, W)
    ref_gacc, ref_illum, ref_sg = z(nb, g.ns, H, W), z(nb, g.ns, g.nz, g.nx), z(nb, g.ns)
    sim._adjoint(sim._geom(nb), 0, g.nt, prep['c2'], prep['sd'], res, hist, m1, m2, ref_gacc, ref_illum, ref_sg,
                 impl=_lib.IMPL_STEPWISE)
    assert float(ref_illum.abs().max()) > 0 and float(ref_sg.abs().max()) > 0
    lib = _lib.load()
    nan = float('nan')
    for tiles in [(0, 0, 0), tiling]:
        state = torch.full((4, nb, g.ns, H, W), nan, device=DEV)
        gacc, illum, sg = torch.full_like(ref_gacc, nan), torch.full_like(ref_illum, nan), torch.full_like(ref_sg, nan)
        arr = _adjoint_group(sim, prep, nb, state, res, hist, gacc, illum, sg, tiles)
        _lib.check(lib.ss_sweep_adjoint_f32(1, arr, 1, _lib.stream_handle(DEV)), 'sweep adjoint')
        torch.cuda.synchronize()
        assert torch.equal(gacc, ref_gacc), f'imaging sum, tiles {tiles}'
        assert torch.equal(illum, ref_illum), f'illumination, tiles {tiles}'
        assert torch.equal(sg, ref_sg), f'source term, tiles {tiles}'
        if tiles == (0, 0, 0):  # one launch from state buffer 0: the final state in buffer 1
            assert torch.equal(state[2], m1) and torch.equal(state[3], m2)
        print(f'  tiles {tiles}: equal', flush=True)


def test_band_wait_bound_fails_into_nan_not_a_hang(monkeypatch):
    """The band kernels' exchange waits are bounded: with the bound forced to zero polls
    from the host (SHOTSWEEP_BAND_SPIN=0) the launch still ends, waits that miss their
    words count as errors (ss_sweep_band_errors) and fill the band's halo with NaN, so
    the gathers show it; the next call with the default bound is exact again (the
    exchange words carry a launch epoch, so nothing of the failed launch is reused)."""
    acq = Acquisition(nz=200, nx=200, nt=60, ns=10)
    case(f'band wait bound forced to zero, {acq.nz}x{acq.nx} nt {acq.nt} x 2 rows')
    sim = gpu.Sim(geom=acq)
    nb, g = 2, acq
    prep = sim._prepare(models(acq, nb, 13))
    H, W = sim.H, sim.W
    ref_out, ref_hist = stepwise_forward(sim, prep)
    lib = _lib.load()
    count = ctypes.c_int(0)

    def errors():
        _lib.check(lib.ss_sweep_band_errors(ctypes.byref(count)), 'band errors')
        return count.value

    def run():
        state = torch.zeros(4, nb, g.ns, H, W, device=DEV)
        hist = torch.empty_like(ref_hist)
        out = torch.empty_like(ref_out)
        arr = _forward_group(sim, prep, nb, state, hist, out)
        _lib.check(lib.ss_sweep_forward_f32(1, arr, 1, _lib.stream_handle(DEV)), 'sweep forward')
        torch.cuda.synchronize()
        return out, hist

    before = errors()
    monkeypatch.setenv('SHOTSWEEP_BAND_SPIN', '0')
    out, _ = run()
    failed = errors() - before
    print(f'  waits past the bound: {failed}, NaN gathers: {int(torch.isnan(out).sum())}', flush=True)
    assert failed > 0 and bool(torch.isnan(out).any())
    monkeypatch.delenv('SHOTSWEEP_BAND_SPIN')
    before = errors()
    out, hist = run()
    assert errors() == before
    assert torch.equal(out, ref_out) and torch.equal(hist, ref_hist)


def test_band_adjoint_wait_bound_fails_into_nan_not_a_hang(monkeypatch):
    """The adjoint band kernel's exchange waits are bounded as the forward's: with the
    bound forced to zero polls from the host the launch ends, the missed waits count
    as errors and the halo's NaN reaches the imaging sum; the next call with the
    default bound equals the stepwise adjoint bit for bit."""
    acq = Acquisition(nz=200, nx=200, nt=61, ns=10)
    case(f'band adjoint wait bound forced to zero, {acq.nz}x{acq.nx} nt {acq.nt} x 2 rows')
    sim = gpu.Sim(geom=acq)
    nb, g = 2, acq
    prep = sim._prepare(models(acq, nb, 14))
    H, W = sim.H, sim.W
    res, hist = stepwise_forward(sim, prep)
    z = lambda *s: torch.zeros(*s, device=DEV)
    m1, m2 = z(nb, g.ns, H, W), z(nb, g.ns, H, W)
    ref_gacc, ref_illum, ref_sg = z(nb, g.ns, H, W), z(nb, g.ns, g.nz, g.nx), z(nb, g.ns)
    sim._adjoint(sim._geom(nb), 0, g.nt, prep['c2'], prep['sd'], res, hist, m1, m2, ref_gacc, ref_illum, ref_sg,
                 impl=_lib.IMPL_STEPWISE)
    lib = _lib.load()
    count = ctypes.c_int(0)

    def errors():
        _lib.check(lib.ss_sweep_band_errors(ctypes.byref(count)), 'band errors')
        return count.value

    def run():
        state = torch.zeros(4, nb, g.ns, H, W, device=DEV)
        gacc, illum, sg = torch.empty_like(ref_gacc), torch.empty_like(ref_illum), torch.empty_like(ref_sg)
        arr = _adjoint_group(sim, prep, nb, state, res, hist, gacc, illum, sg)
        _lib.check(lib.ss_sweep_adjoint_f32(1, arr, 1, _lib.stream_handle(DEV)), 'sweep adjoint')
        torch.cuda.synchronize()
        return gacc, illum, sg

    before = errors()
    monkeypatch.setenv('SHOTSWEEP_BAND_SPIN', '0')
    gacc, _, _ = run()
    failed = errors() - before
    print(f'  waits past the bound: {failed}, NaN imaging-sum cells: {int(torch.isnan(gacc).sum())}', flush=True)
    assert failed > 0 and bool(torch.isnan(gacc).any())
    monkeypatch.delenv('SHOTSWEEP_BAND_SPIN')
    before = errors()
    gacc, illum, sg = run()
    assert errors() == before
    assert torch.equal(gacc, ref_gacc) and torch.equal(illum, ref_illum) and torch.equal(sg, ref_sg)


def test_captured_sweep_keeps_to_tiled_kernels():
    """A sweep call captured into a graph on a band grid (200 x 200 models, pad 40)
    runs the tiled kernels, since a captured cooperative launch would lose its
    co-residency guarantee: three replays, with eager work and synchronisations in
    between, each give the gathers and history of the eager band call bit for bit."""
    acq = Acquisition(nz=200, nx=200, nt=60, ns=10)
    case(f'captured sweep {acq.nz}x{acq.nx} nt {acq.nt} x 1 row')
    sim = gpu.Sim(geom=acq)
    nb, g = 1, acq
    prep = sim._prepare(models(acq, nb, 14))
    H, W = sim.H, sim.W
    lib = _lib.load()
    state = torch.zeros(4, nb, g.ns, H, W, device=DEV)
    ref_hist = torch.empty(nb, g.ns, g.nt, H, W, device=DEV)
    ref_out = torch.empty(nb, g.ns, g.nt, g.nx, device=DEV)
    arr = _forward_group(sim, prep, nb, state, ref_hist, ref_out)
    _lib.check(lib.ss_sweep_forward_f32(1, arr, 0, _lib.stream_handle(DEV)), 'eager sweep')
    torch.cuda.synchronize()
    hist, out = torch.full_like(ref_hist, float('nan')), torch.full_like(ref_out, float('nan'))
    arr = _forward_group(sim, prep, nb, state, hist, out)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        _lib.check(lib.ss_sweep_forward_f32(1, arr, 0, _lib.stream_handle(DEV)), 'captured sweep')
    for rep in range(3):
        hist.fill_(float('nan'))
        out.fill_(float('nan'))
        x = torch.randn(1000, device=DEV)
        _ = float((x * x).sum())      # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref_out) and torch.equal(hist, ref_hist), f'replay {rep}'
