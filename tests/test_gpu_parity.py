"""GPU parity of the HIP path (through the C ABI) with the reference's golden
vectors and the CPU oracle.

Bar: gathers, illumination and the stencil operators are bit-identical to the
reference FP32/FP64 CPU path; gradient and image pixels off the padding fold's
edge strips are bit-identical, and the edge pixels (whose fold sums run in a
different order) stay within the rounding bound of their sums; whole gradients
and images are also checked within 1e-5 relative L2 in FP32 and 1e-12 in FP64.
"""
import numpy as np
import pytest
import torch

from oracle import fd
from shotsweep_amd import _lib, gpu, synth
from shotsweep_amd.constants import Acquisition

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda')
F32_TOL = 1e-5
F64_TOL = 1e-12
SMALL = Acquisition(nz=24, nx=28, nt=260, ns=3)
ORACLE_SMALL = fd.Geometry(nz=24, nx=28, nt=260, ns=3)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(name):
    print(f'\n[case] {name}', flush=True)


@pytest.fixture(autouse=True)
def sync_after():
    yield
    torch.cuda.synchronize()


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_stencil_operators_match_oracle(dtype):
    case(f'lap/lap_adjoint/fold {dtype}')
    g = torch.Generator().manual_seed(8)
    x = torch.randn(2, 3, 19, 23, generator=g, dtype=dtype)
    np.testing.assert_array_equal(gpu.lap(x.to(DEV)).cpu().numpy(), fd.lap(x).numpy())
    np.testing.assert_array_equal(gpu.lap_adjoint(x.to(DEV)).cpu().numpy(), fd.lap_adjoint(x).numpy())
    y = torch.randn(2, 3, 19, 23, generator=g, dtype=dtype)
    lhs = float((gpu.lap(x.to(DEV)) * y.to(DEV)).sum())
    rhs = float((x.to(DEV) * gpu.lap_adjoint(y.to(DEV))).sum())
    assert abs(lhs - rhs) <= (1e-4 if dtype == torch.float32 else 1e-10) * abs(lhs)
    f = gpu.fold_edge_padding(x.to(DEV), 5).cpu().numpy()
    np.testing.assert_allclose(f, fd.fold_edge_padding(x, 5, 9, 13).numpy(),
                               rtol=1e-6 if dtype == torch.float32 else 1e-14, atol=1e-5)


@pytest.mark.parametrize('impl,dtype,tag', [(_lib.IMPL_STEPWISE, torch.float32, 'f32'),
                                            (_lib.IMPL_STEPWISE, torch.float64, 'f64'),
                                            (_lib.IMPL_RESIDENT, torch.float32, 'f32')])
def test_gathers_bit_identical_to_reference(golden, impl, dtype, tag):
    case(f'gathers impl={impl} {tag}')
    sim = gpu.Sim(device='cuda', dtype=dtype, impl=impl)
    obs = sim.gathers(golden['true'])
    np.testing.assert_array_equal(obs, golden[f'obs_{tag}'])


@pytest.mark.parametrize('impl,dtype,tag,history', [
    (_lib.IMPL_RESIDENT, torch.float32, 'f32', 'full'),
    (_lib.IMPL_RESIDENT, torch.float32, 'f32', 'checkpoint'),
    (_lib.IMPL_STEPWISE, torch.float32, 'f32', 'full'),
    (_lib.IMPL_STEPWISE, torch.float64, 'f64', 'checkpoint'),
])
def test_images_match_reference(golden, impl, dtype, tag, history):
    case(f'images impl={impl} {tag} history={history}')
    sim = gpu.Sim(device='cuda', dtype=dtype, impl=impl, history=history)
    obs = golden[f'obs_{tag}']
    img, illum = sim.images(golden['backgrounds'], np.repeat(obs, 3, 0), v_mute=golden['v_mute'])
    np.testing.assert_array_equal(illum, golden[f'illum_{tag}'])
    tol = F32_TOL if tag == 'f32' else F64_TOL
    e = rel(img, golden[f'img_{tag}'])
    print(f'  image rel L2 vs reference {tag}: {e:.3e}; vs reference FP64: {rel(img, golden["img_f64"]):.3e}')
    assert e < tol


@pytest.mark.parametrize('impl,dtype,tag', [(_lib.IMPL_RESIDENT, torch.float32, 'f32'),
                                            (_lib.IMPL_STEPWISE, torch.float64, 'f64')])
def test_gradient_matches_reference(golden, impl, dtype, tag):
    case(f'gradient impl={impl} {tag}')
    sim = gpu.Sim(device='cuda', dtype=dtype, impl=impl)
    grad, illum = sim.gradient(golden['backgrounds'][:1], golden[f'obs_{tag}'])
    np.testing.assert_array_equal(illum, golden[f'grad_illum_{tag}'])
    e = rel(grad, golden[f'grad_{tag}'])
    print(f'  gradient rel L2 vs reference {tag}: {e:.3e}')
    assert e < (F32_TOL if tag == 'f32' else F64_TOL)


def _fold_terms(pad, nz, nx):
    """Terms each model pixel of fold_edge_padding sums: 1 own value, pad per padding
    strip, pad^2 per corner block (gpu.py:56-67)."""
    n = np.ones((nz, nx))
    n[0, :] += pad
    n[-1, :] += pad
    n[:, 0] += pad
    n[:, -1] += pad
    for i in (0, -1):
        for j in (0, -1):
            n[i, j] += pad * pad
    return n


def test_gradient_interior_bit_identical_edges_within_fold_bound(golden):
    """FP32 gradient against the reference's: every pixel off the model's edge rows and
    columns (including the source pixels of the three inner shots, with their
    source-amplitude term) is the reference's to the bit; only the fold's padding-strip
    and corner-block sums run in another order, so each edge pixel is within the
    summation bound 2 n eps sum|terms| of that pixel (n terms; sum|terms| folded from
    the oracle's unfolded imaging sum)."""
    case('gradient interior bitwise, edge pixels within the fold rounding bound')
    bg, obs = golden['backgrounds'][:1], golden['obs_f32']
    grad, illum = gpu.Sim(device='cuda').gradient(bg, obs)
    want = golden['grad_f32']
    inner = np.zeros(want.shape, bool)
    inner[..., 1:-1, 1:-1] = True
    np.testing.assert_array_equal(grad[inner], want[inner])
    osim = fd.OracleSim()
    og, _ = osim.gradient(bg, obs)
    np.testing.assert_array_equal(og, want)          # the oracle is the reference here
    mag = fd.fold_edge_padding(osim.last_scaled.abs(), 40, 70, 70).numpy().astype(np.float64)
    bound = 2 * _fold_terms(40, 70, 70) * 2.0 ** -24 * mag + 2.0 ** -23 * np.abs(want)
    err = np.abs(grad.astype(np.float64) - want)
    worst = float((err / bound)[~inner].max())
    print(f'  edge pixels: worst error / bound {worst:.3f}; interior bit-identical')
    assert np.all(err[~inner] <= bound[~inner])


def test_images_interior_bit_identical(golden):
    """FP32 images without the mute against the oracle (the reference restated, pinned
    to its golden outputs): the Laplacian filter of a pixel reads pixels up to 2 away,
    so every pixel at least 3 from the model's edge is the oracle's to the bit; the
    rest within 1e-5 relative L2.  (With the direct-wave mute the weights' cosine is
    correctly rounded here and torch's vectorised CPU cosine in the reference, up to
    4 ulp apart, so muted images are compared by relative L2 only.)"""
    case('images interior bitwise, no mute')
    obs = golden['obs_f32']
    bgs = golden['backgrounds'][:2]
    img, illum = gpu.Sim(device='cuda').images(bgs, np.repeat(obs, 2, 0))
    want, want_illum = fd.OracleSim().images(bgs, np.repeat(obs, 2, 0))
    np.testing.assert_array_equal(illum, want_illum)
    inner = np.zeros(want.shape, bool)
    inner[..., 3:-3, 3:-3] = True
    np.testing.assert_array_equal(img[inner], want[inner])
    e = rel(img[~inner], want[~inner])
    print(f'  outer pixels rel L2 {e:.2e}; interior bit-identical')
    assert e < F32_TOL


def test_checkpoint_recompute_equals_full_history(golden):
    case('checkpoint vs full history bitwise')
    obs = np.repeat(golden['obs_f32'], 3, 0)
    full = gpu.Sim(device='cuda', history='full').images(golden['backgrounds'], obs, v_mute=golden['v_mute'])
    # resident segments (recompute overlapped on a side stream) and tiled segments (the sweep
    # kernels over each segment's step range, resuming state): a second call on the same Sim
    # (reused buffers, events, graphs) and an interval of 1000 (one segment) agree too
    for seg_impl in ('resident', 'tiled', 'auto'):
        sim = gpu.Sim(device='cuda', history='checkpoint', checkpoint_interval=73, segment_impl=seg_impl)
        ck = sim.images(golden['backgrounds'], obs, v_mute=golden['v_mute'])
        again = sim.images(golden['backgrounds'], obs, v_mute=golden['v_mute'])
        one = gpu.Sim(device='cuda', history='checkpoint', checkpoint_interval=1000, segment_impl=seg_impl).images(
            golden['backgrounds'], obs, v_mute=golden['v_mute'])
        for other in (ck, again, one):
            np.testing.assert_array_equal(full[0], other[0], err_msg=seg_impl)
            np.testing.assert_array_equal(full[1], other[1], err_msg=seg_impl)


@pytest.mark.parametrize('interval', [1, 2, 7])
def test_short_resident_segments_equal_full_history(golden, interval):
    """Resident checkpoint segments of 1, 2 and 7 steps (the paired adjoint's history
    ring then wraps into the next segment after every few pairs or at once) give the
    full-history images and illumination bit for bit."""
    case(f'resident segments of {interval} steps vs full history')
    obs = np.repeat(golden['obs_f32'], 3, 0)
    full = gpu.Sim(device='cuda', history='full').images(golden['backgrounds'], obs, v_mute=golden['v_mute'])
    ck = gpu.Sim(device='cuda', history='checkpoint', checkpoint_interval=interval, segment_impl='resident').images(
        golden['backgrounds'], obs, v_mute=golden['v_mute'])
    np.testing.assert_array_equal(full[0], ck[0])
    np.testing.assert_array_equal(full[1], ck[1])


@pytest.mark.parametrize('nx', [28, 27])
def test_tiled_segments_equal_resident_segments(nx):
    """Checkpointed gradients on a small grid (paired kernels for an even width, cell
    kernels for an odd one): tiled segments, resident segments and the whole history
    agree bit for bit, with a short last segment (nt 260, interval 40)."""
    case(f'tiled vs resident checkpoint segments 24x{nx} nt=260 ns=3 interval 40')
    geom = Acquisition(nz=24, nx=nx, nt=260, ns=3)
    rng = np.random.default_rng(5)
    v = 1800 + 900 * rng.random((2, 24, nx))
    obs = gpu.Sim(device='cuda', pad=12, geom=geom).gathers(v) * 0.9
    vm = np.array([1900.0, 2100.0])
    want = gpu.Sim(device='cuda', pad=12, geom=geom, history='full').gradient(v, obs, v_mute=vm)
    for seg_impl in ('tiled', 'resident'):
        sim = gpu.Sim(device='cuda', pad=12, geom=geom, history='checkpoint', checkpoint_interval=40,
                      segment_impl=seg_impl)
        for _ in range(2):
            got = sim.gradient(v, obs, v_mute=vm)
            np.testing.assert_array_equal(got[0], want[0], err_msg=seg_impl)
            np.testing.assert_array_equal(got[1], want[1], err_msg=seg_impl)


@pytest.mark.parametrize('nx', [28, 27])
def test_tiled_checkpoint_forward_equals_resident(nx):
    """The checkpointed forward of small batches on the tiled kernels (state carried
    from segment to segment, no Laplacians kept) records the same gathers as the
    resident forward, and the adjoint from either tape (forward and adjoint as two
    native calls, as autograd runs them) gives the same gradient, bit for bit, with
    a short last segment."""
    case(f'tiled vs resident checkpointed forward 24x{nx} nt=260 ns=3 interval 40')
    geom = Acquisition(nz=24, nx=nx, nt=260, ns=3)
    v = 1800 + 900 * np.random.default_rng(6).random((2, 24, nx))
    seed = torch.tensor(np.random.default_rng(7).standard_normal((2, 3, 260, nx)), dtype=torch.float32, device=DEV)
    out = {}
    for seg_impl in ('tiled', 'resident'):
        sim = gpu.Sim(device='cuda', pad=12, geom=geom, history='checkpoint', checkpoint_interval=40,
                      segment_impl=seg_impl)
        prep = sim._prepare(v)
        gathers, tape = sim._forward_tape(prep, segment=40, fresh=True)
        grad, illum = sim._adjoint_tape(prep, tape, seed)
        out[seg_impl] = (gathers.cpu().numpy(), grad.cpu().numpy(), illum.cpu().numpy())
    for a, b in zip(out['tiled'], out['resident']):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_mute_weights_match_oracle(dtype):
    case(f'mute weights {dtype}')
    vm = np.array([1550.0, 2000.0, 3100.0])
    got = gpu.Sim(device='cuda', dtype=dtype).mute_weights(vm).cpu().numpy()
    want = fd.OracleSim(dtype=dtype).mute_weights(vm).numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=4 * np.finfo(np.float32 if dtype == torch.float32
                                                                      else np.float64).eps)
    # before and after the 30 ms taper the weights are exactly 0 and 1, as in the oracle
    ends = (want == 0) | (want == 1)
    assert ends.mean() > 0.9
    np.testing.assert_array_equal(got[ends], want[ends])


def test_calibrate_matches_reference(golden):
    case('calibrate')
    cal_obs = 0.7 * np.roll(golden['obs_f32'].astype(np.float64), 3, axis=2)
    np.testing.assert_array_equal(gpu.calibrate(golden['true'], cal_obs), golden['calibrated'])


def test_mixed_models_resident_vs_oracle():
    case('mixed models 2x3 backgrounds resident vs oracle FP32')
    models = synth.mixed(2, seed=11)
    sim = gpu.Sim(device='cuda')
    osim = fd.OracleSim()
    obs = sim.gathers(models)
    np.testing.assert_array_equal(obs, osim.gathers(models))
    np.testing.assert_array_equal(sim.gathers(models, chunk=1), obs)   # chunking (gpu.py:138-140)
    trend = np.linspace(0, 1500, 70)
    bgs = np.concatenate([np.stack([np.tile(np.clip(m[0].mean() + r * trend, 1500, 4500)[:, None], (1, 70))
                                    for r in (0.6, 1.0, 1.4)]) for m in models])
    vm = np.repeat(models[:, 0].mean(1), 3)
    img, illum = sim.images(bgs, np.repeat(obs, 3, 0), v_mute=vm)
    oimg, oillum = osim.images(bgs, np.repeat(obs, 3, 0), v_mute=vm)
    np.testing.assert_array_equal(illum, oillum)
    e = rel(img, oimg)
    print(f'  image rel L2 vs oracle: {e:.3e}')
    assert e < F32_TOL


def test_rectangular_geometry_vs_oracle():
    case('rectangular 24x28 nt=260 ns=3 stepwise+resident vs oracle')
    rng = np.random.default_rng(2)
    v = 1800 + 900 * rng.random((2, 24, 28))
    want = fd.OracleSim(pad=12, geom=ORACLE_SMALL).gathers(v)
    for impl in (_lib.IMPL_STEPWISE, _lib.IMPL_RESIDENT):
        got = gpu.Sim(device='cuda', pad=12, geom=SMALL, impl=impl).gathers(v)
        np.testing.assert_array_equal(got, want)
    obs = want * 0.9
    og, oi = fd.OracleSim(pad=12, geom=ORACLE_SMALL).gradient(v, obs, v_mute=np.array([1900.0, 2100.0]))
    gg, gi = gpu.Sim(device='cuda', pad=12, geom=SMALL).gradient(v, obs, v_mute=np.array([1900.0, 2100.0]))
    np.testing.assert_array_equal(gi, oi)
    assert rel(gg, og) < F32_TOL


@pytest.mark.parametrize('nz,nx,ns', [(2, 2, 2), (2, 5, 1), (6, 3, 3), (8, 40, 32)])
def test_edge_geometries_vs_oracle(nz, nx, ns):
    """Tiny models (receiver row 1 of 2; one padding fold pixel gets top and bottom
    strips, or several corner blocks), a single shot, and the maximum of 32 shots
    with sources sharing columns: gathers and illumination bit-identical to the
    oracle, gradients within 1e-5."""
    case(f'edge geometry {nz}x{nx} ns={ns} nt=120 pad 6 vs oracle')
    geom = Acquisition(nz=nz, nx=nx, nt=120, ns=ns)
    ogeom = fd.Geometry(nz=nz, nx=nx, nt=120, ns=ns, src=tuple(geom.src))
    v = 1800 + 900 * np.random.default_rng(nz * 100 + nx).random((2, nz, nx))
    vm = np.array([1900.0, 2100.0])
    want = fd.OracleSim(pad=6, geom=ogeom).gathers(v)
    np.testing.assert_array_equal(gpu.Sim(device='cuda', pad=6, geom=geom).gathers(v), want)
    obs = want * 0.9
    # mute start 20 ms, so that the 120-step traces keep unmuted samples and the taper
    og, oi = fd.OracleSim(pad=6, geom=ogeom).gradient(v, obs, v_mute=vm, t_mute=0.02)
    gg, gi = gpu.Sim(device='cuda', pad=6, geom=geom).gradient(v, obs, v_mute=vm, t_mute=0.02)
    assert np.abs(og).max() > 0
    np.testing.assert_array_equal(gi, oi)
    assert rel(gg, og) < F32_TOL


@pytest.mark.parametrize('laplacian,eps,mute', [(False, 1e-3, False), (True, 1e-2, True), (False, 1e-4, True)])
def test_image_options_vs_oracle(laplacian, eps, mute):
    """images() with and without the Laplacian filter, other illumination
    regularisations and without a mute (gpu.py:196-210): illumination bit-identical
    to the oracle, images within 1e-5."""
    case(f'image options laplacian={laplacian} eps={eps} mute={mute} 24x28 nt=260')
    rng = np.random.default_rng(8)
    v = 1800 + 900 * rng.random((2, 24, 28))
    obs = fd.OracleSim(pad=12, geom=ORACLE_SMALL).gathers(v) * 0.9
    vm = np.array([1900.0, 2100.0]) if mute else None
    oi, oil = fd.OracleSim(pad=12, geom=ORACLE_SMALL).images(v, obs, v_mute=vm, t_mute=0.05, eps=eps,
                                                               laplacian=laplacian)
    gi, gil = gpu.Sim(device='cuda', pad=12, geom=SMALL).images(v, obs, v_mute=vm, t_mute=0.05, eps=eps,
                                                                laplacian=laplacian)
    np.testing.assert_array_equal(gil, oil)
    assert np.abs(oi).max() > 0 and rel(gi, oi) < F32_TOL


def test_forward_checkpoints_vs_oracle():
    """forward(v, store=True) (gpu.py:117-136): gathers, the (p0, p1) checkpoints
    before every checkpoint_interval-th step (interval 73, a short last one) and
    c2 bit-identical to the oracle's."""
    case('forward store=True checkpoints 24x28 nt=260 interval 73 vs oracle')
    v = 1800 + 900 * np.random.default_rng(9).random((2, 24, 28))
    got = gpu.Sim(device='cuda', pad=12, geom=SMALL, checkpoint_interval=73).forward(v, store=True)
    want = fd.OracleSim(pad=12, geom=ORACLE_SMALL, checkpoint_interval=73).forward(v, store=True)
    np.testing.assert_array_equal(got[0].cpu().numpy(), want[0].numpy())
    assert sorted(got[1]) == sorted(want[1]) == [0, 73, 146, 219]
    for it, (p0, p1) in want[1].items():
        np.testing.assert_array_equal(got[1][it][0].cpu().numpy(), p0.numpy(), err_msg=f'p0 at {it}')
        np.testing.assert_array_equal(got[1][it][1].cpu().numpy(), p1.numpy(), err_msg=f'p1 at {it}')
    np.testing.assert_array_equal(got[2].cpu().numpy().reshape(want[2].shape), want[2].numpy())


def test_dot_product_fp64():
    case('dot product <y,Fx> = <F*y,x> FP64')
    sim = gpu.Sim(device='cuda', dtype=torch.float64, pad=12, geom=SMALL)
    rng = np.random.default_rng(5)
    v = 2000 + 300 * rng.random((2, SMALL.nz, SMALL.nx))
    x = rng.standard_normal(v.shape)
    y = torch.tensor(rng.standard_normal((2, SMALL.ns, SMALL.nt, SMALL.nx)), device=DEV)
    Fx = sim.born(v, x)
    grad, _, _ = sim._gradient_device(sim._prepare(v), res=y)
    lhs = float((y * Fx).sum())
    rhs = float((grad.sum(1).cpu().numpy() * x).sum())
    print(f'  dot-product relative mismatch {abs(lhs - rhs) / abs(lhs):.3e}')
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_gradient_finite_differences_fp64():
    case('finite differences FP64')
    sim = gpu.Sim(device='cuda', dtype=torch.float64, pad=12, geom=SMALL)
    v = np.full((1, SMALL.nz, SMALL.nx), 2000.0)
    truth = v.copy()
    truth[:, 8:] = 2400.0
    obs = sim.gathers(truth)
    grad = sim.gradient(v, obs)[0].sum(1)
    rng = np.random.default_rng(9)
    for k in range(4):
        d = np.zeros_like(v)
        d[0, rng.integers(0, SMALL.nz), rng.integers(0, SMALL.nx)] = 1

        def loss(m):
            r = sim.gathers(m) - obs
            return 0.5 * np.sum(r * r)
        eps = 0.1
        num = (loss(v + eps * d) - loss(v - eps * d)) / (2 * eps)
        assert abs(num - np.sum(grad * d)) / max(abs(num), 1.0) < 1e-4


def test_errors_match_reference():
    case('validation errors')
    sim = gpu.Sim(device='cuda')
    with pytest.raises(ValueError, match='velocities must be finite positive'):
        sim.gathers(np.full((1, 70, 69), 2000.0))
    with pytest.raises(ValueError, match='observations must be finite'):
        sim.gradient(np.full((1, 70, 70), 2000.0), np.zeros((1, 5, 999, 70)))
    with pytest.raises(ValueError, match='eps must be positive'):
        sim.images(np.full((1, 70, 70), 2000.0), np.zeros((1, 5, 1000, 70)), eps=0)
    with pytest.raises(ValueError, match='mute velocities'):
        sim.mute_weights(np.array([-1.0]))
    with pytest.raises(ValueError, match='wavelet must contain'):
        gpu.Sim(np.zeros(1000), device='cuda')
    with pytest.raises(ValueError, match='unsupported device'):
        gpu.Sim(device='xla')


def test_wide_shallow_grid_records_every_receiver():
    """A grid wider than a workgroup (2 x 1100 model, pad 2: 1100 receivers for 1024
    threads): gathers bit-identical to the oracle on the resident and stepwise
    forwards, and the checkpointed gradient (tiled segments, tiles at most 1024
    columns wide) equal to the oracle's illumination bit for bit, gradient within 1e-5."""
    case('wide shallow grid 2x1100 pad 2 nt 80 ns 2 vs oracle')
    geom = Acquisition(nz=2, nx=1100, nt=80, ns=2)
    ogeom = fd.Geometry(nz=2, nx=1100, nt=80, ns=2, src=tuple(geom.src))
    v = 1800 + 900 * np.random.default_rng(12).random((1, 2, 1100))
    want = fd.OracleSim(pad=2, geom=ogeom).gathers(v)
    assert np.abs(want[..., 1024:]).max() > 0
    for impl in (_lib.IMPL_RESIDENT, _lib.IMPL_STEPWISE):
        np.testing.assert_array_equal(gpu.Sim(device='cuda', pad=2, geom=geom, impl=impl).gathers(v), want)
    obs = want * 0.9
    og, oi = fd.OracleSim(pad=2, geom=ogeom, checkpoint_interval=30).gradient(v, obs, t_mute=0.0,
                                                                             v_mute=np.array([5000.0]))
    gg, gi = gpu.Sim(device='cuda', pad=2, geom=geom, history='checkpoint', checkpoint_interval=30,
                     segment_impl='tiled').gradient(v, obs, v_mute=np.array([5000.0]), t_mute=0.0)
    np.testing.assert_array_equal(gi, oi)
    assert rel(gg, og) < F32_TOL


@pytest.mark.parametrize('history', ['full', 'checkpoint'])
def test_pinned_host_observations_equal_device_observations(history):
    """images_device with pinned host observations (uploaded on a copy stream beside
    the forward stage, the adjoint stage after the upload) gives the same bits as
    with device observations, for the whole history and checkpoint segments."""
    case(f'pinned host observations, history {history}')
    sim = gpu.Sim(device=DEV, history=history)
    models = synth.mixed(2, seed=31)
    obs = sim.forward(models)[0].contiguous()
    bgs = np.repeat(synth.smooth(models, 6.0), 3, axis=0)
    vm = torch.full((6,), 1900.0, device=DEV)
    prep = sim._prepare(bgs)
    img_d, ill_d = (t.cpu() for t in sim.images_device(prep, obs, 3, vm))
    pinned = obs.cpu().pin_memory()
    img_h, ill_h = (t.cpu() for t in sim.images_device(prep, pinned, 3, vm))
    assert torch.equal(img_h, img_d) and torch.equal(ill_h, ill_d)
    with pytest.raises(ValueError, match='pinned'):
        sim.images_device(prep, obs.cpu(), 3, vm)


def test_public_resident_adjoint_equals_stepwise():
    """The adjoint entry point (ss_adjoint_f32) on the 150 x 150 plane with the
    resident kernels and no c2*A plane (the column-strip adjoint then forms c2*A
    from its sponge table) equals the stepwise kernels bit for bit: imaging sum,
    illumination, source term and final state, over the whole record and over two
    segments of odd length that carry the state between calls (gpu.py:161-184)."""
    case('public resident adjoint vs stepwise, whole record and two segments')
    sim = gpu.Sim(device='cuda')
    g = sim.geom
    models = synth.mixed(1, seed=11)
    prep = sim._prepare(np.repeat(models, 2, 0))
    nb, H, W = prep['nb'], sim.H, sim.W
    geom = sim._geom(nb)
    z = lambda *s: torch.zeros(*s, device=DEV)
    p0, p1 = z(nb, g.ns, H, W), z(nb, g.ns, H, W)
    hist = torch.empty(nb, g.ns, g.nt, H, W, device=DEV)
    sim._forward(geom, 0, g.nt, prep['c2'], prep['amp'], p0, p1, hist=hist, impl=_lib.IMPL_STEPWISE)
    gen = torch.Generator(device=DEV).manual_seed(5)
    res = torch.randn(nb, g.ns, g.nt, g.nx, generator=gen, device=DEV)

    def run(impl, cuts):
        m1, m2, gacc = z(nb, g.ns, H, W), z(nb, g.ns, H, W), z(nb, g.ns, H, W)
        illum, sg = z(nb, g.ns, g.nz, g.nx), z(nb, g.ns)
        for t0, t1 in cuts:  # reverse time: later segments first
            seg = hist[:, :, t0:t1].contiguous()
            sim._adjoint(geom, t0, t1, prep['c2'], prep['sd'], res, seg, m1, m2, gacc, illum, sg, impl=impl)
        return m1, m2, gacc, illum, sg

    want = run(_lib.IMPL_STEPWISE, [(0, g.nt)])
    for cuts in ([(0, g.nt)], [(377, g.nt), (0, 377)]):
        got = run(_lib.IMPL_RESIDENT, cuts)
        for name, a, b in zip(('m1', 'm2', 'imaging sum', 'illumination', 'source term'), got, want):
            assert torch.equal(a, b), f'{name}, segments {cuts}'


def test_public_resident_forward_equals_stepwise():
    """The forward entry point (ss_forward_f32) on the 150 x 150 plane with the resident
    kernels equals the stepwise kernels bit for bit: gathers, the Laplacian history
    including its zero two-cell rim (gpu.py:35), and the final state, over the whole
    record and over two segments of odd length (gpu.py:111-136)."""
    case('public resident forward vs stepwise, whole record and two segments')
    sim = gpu.Sim(device='cuda')
    g = sim.geom
    prep = sim._prepare(np.repeat(synth.mixed(1, seed=12), 2, 0))
    nb, H, W = prep['nb'], sim.H, sim.W
    geom = sim._geom(nb)

    def run(impl, cuts):
        p0 = torch.zeros(nb, g.ns, H, W, device=DEV)
        p1 = torch.zeros_like(p0)
        out = torch.empty(nb, g.ns, g.nt, g.nx, device=DEV)
        hists = []
        for t0, t1 in cuts:
            h = torch.empty(nb, g.ns, t1 - t0, H, W, device=DEV)
            sim._forward(geom, t0, t1, prep['c2'], prep['amp'], p0, p1, gathers=out, hist=h, impl=impl)
            hists.append(h)
        return out, torch.cat(hists, 2), p0, p1

    want = run(_lib.IMPL_STEPWISE, [(0, g.nt)])
    assert float(want[1][..., :2, :].abs().max()) == 0.0 and float(want[1][..., :, -2:].abs().max()) == 0.0
    for cuts in ([(0, g.nt)], [(0, 377), (377, g.nt)]):
        got = run(_lib.IMPL_RESIDENT, cuts)
        for name, a, b in zip(('gathers', 'history', 'p0', 'p1'), got, want):
            assert torch.equal(a, b), f'{name}, segments {cuts}'
