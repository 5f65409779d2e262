"""CPU pins of the oracle (oracle/fd.py) against the reference's golden vectors,
closed forms, brute force and the invariants of the discrete operators."""
import numpy as np
import pytest
import torch

from oracle import fd

SMALL = fd.Geometry(nz=24, nx=28, nt=260, ns=3)   # fast FP64 checks on a rectangular grid


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope='module', autouse=True)
def threads():
    torch.set_num_threads(4)


# ----------------------------------------------------------- golden vectors
@pytest.mark.parametrize('tag,dtype', [('f32', torch.float32), ('f64', torch.float64)])
def test_matches_reference_golden_bit_for_bit(golden, tag, dtype):
    sim = fd.OracleSim(dtype=dtype)
    obs = sim.gathers(golden['true'])
    np.testing.assert_array_equal(obs, golden[f'obs_{tag}'])
    img, illum = sim.images(golden['backgrounds'], np.repeat(obs, 3, 0), v_mute=golden['v_mute'])
    np.testing.assert_array_equal(img, golden[f'img_{tag}'])
    np.testing.assert_array_equal(illum, golden[f'illum_{tag}'])
    grad, gillum = sim.gradient(golden['backgrounds'][:1], obs)
    np.testing.assert_array_equal(grad, golden[f'grad_{tag}'])
    np.testing.assert_array_equal(gillum, golden[f'grad_illum_{tag}'])


def test_calibrate_matches_reference(golden):
    cal_obs = 0.7 * np.roll(golden['obs_f32'].astype(np.float64), 3, axis=2)
    np.testing.assert_array_equal(fd.calibrate(golden['true'], cal_obs), golden['calibrated'])


# ------------------------------------------------------------- closed forms
def test_constants_and_sources():
    assert fd.REFERENCE.src == (0, 17, 34, 52, 69)
    assert fd.even_sources(140, 7) == [0, 23, 46, 70, 93, 116, 139]
    w = fd.RICKER
    assert w.shape == (1000,) and np.argmax(w) == 80 and w[80] == 1.0
    # zero crossings of (1-2a)e^-a where a = 1/2: t = 1.2/f0 +- 1/(sqrt(2) pi f0)
    tz = 1.2 / 15 + 1 / (np.sqrt(2) * np.pi * 15)
    i = int(tz / 1e-3)
    assert w[i] > 0 > w[i + 1]


def test_damping_closed_form():
    A, B = fd.damping(fd.REFERENCE, 40, 1e-5)
    assert A.shape == (150, 150)
    assert np.all(A[40:110, 40:110] == 1) and np.all(B[40:110, 40:110] == 1)
    a = 3 * 4500 / (2 * 40 * 10) * np.log(1e5) * 2 * 1e-3 / 2
    assert A[0, 0] == pytest.approx(1 / (1 + a), rel=1e-15) and B[-1, -1] == pytest.approx(1 - a, rel=1e-15)
    a1 = 3 * 4500 / (2 * 40 * 10) * np.log(1e5) * (20 / 40) ** 2 * 1e-3 / 2
    assert A[20, 75] == pytest.approx(1 / (1 + a1), rel=1e-15)
    assert np.array_equal(A, A[::-1]) and np.array_equal(A, A.T)


def _poly(z, x):
    return 0.3 * z ** 3 - 1.1 * z ** 2 * x + 0.7 * x ** 4 + 2.0 * z * x - 5.0, \
        1.8 * z - 2.2 * x + 8.4 * x ** 2


def test_lap_exact_on_polynomials_and_zero_border():
    z, x = np.meshgrid(np.arange(13.0), np.arange(11.0), indexing='ij')
    f, d2 = _poly(z, x)
    out = fd.lap(torch.tensor(f)[None]).numpy()[0]
    np.testing.assert_allclose(out[2:-2, 2:-2], d2[2:-2, 2:-2], rtol=1e-12, atol=1e-10)
    assert np.all(out[:2] == 0) and np.all(out[-2:] == 0) and np.all(out[:, :2] == 0) and np.all(out[:, -2:] == 0)


def test_planted_tap_error_is_caught():
    """A wrong weight in the oracle's stencil fails the closed form above."""
    z, x = np.meshgrid(np.arange(13.0), np.arange(11.0), indexing='ij')
    f, d2 = _poly(z, x)
    p = torch.tensor(f)[None]
    mid = p[..., 2:-2, 2:-2]
    wrong = (-p[..., :-4, 2:-2] + 16 * p[..., 1:-3, 2:-2] - 30 * mid + 16 * p[..., 3:-1, 2:-2] - p[..., 4:, 2:-2]
             - p[..., 2:-2, :-4] + 15 * p[..., 2:-2, 1:-3] - 30 * mid + 16 * p[..., 2:-2, 3:-1] - p[..., 2:-2, 4:]) / 12
    assert not np.allclose(wrong.numpy()[0], d2[2:-2, 2:-2], rtol=1e-6)


def test_lap_adjoint_is_transpose_including_boundary():
    g = torch.Generator().manual_seed(8)
    x = torch.randn(2, 9, 11, generator=g, dtype=torch.float64)
    y = torch.randn(2, 9, 11, generator=g, dtype=torch.float64)
    assert float((fd.lap(x) * y).sum()) == pytest.approx(float((x * fd.lap_adjoint(y)).sum()), rel=1e-13)


def test_lap_adjoint_brute_force_matrix():
    H, W = 7, 8
    eye = torch.eye(H * W, dtype=torch.float64).view(H * W, H, W)
    M = fd.lap(eye).reshape(H * W, H * W)          # row k = lap(e_k)  => M = lap^T as a matrix on vec
    y = torch.randn(H, W, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    expected = (M @ y.reshape(-1)).view(H, W)       # lap^T y
    np.testing.assert_allclose(fd.lap_adjoint(y[None])[0].numpy(), expected.numpy(), rtol=1e-14, atol=1e-14)


def test_fold_is_transpose_of_edge_padding():
    rng = np.random.default_rng(3)
    nz, nx, p = 6, 9, 4
    g = rng.standard_normal((2, 3, nz + 2 * p, nx + 2 * p))
    v = rng.standard_normal((2, 3, nz, nx))
    lhs = (fd.fold_edge_padding(torch.tensor(g), p, nz, nx).numpy() * v).sum()
    rhs = (g * np.pad(v, ((0, 0), (0, 0), (p, p), (p, p)), mode='edge')).sum()
    assert lhs == pytest.approx(rhs, rel=1e-13)


def test_mute_weights_closed_form():
    sim = fd.OracleSim(dtype=torch.float64)
    w = sim.mute_weights(np.array([2000.0]), t_mute=0.12).numpy()[0]
    # shot 2 (x = 340 m), receiver 44 (x = 440 m): h = 100 m, t0 = 0.05 + 0.12 = 0.17 s
    t0 = 100 / 2000 + 0.12
    assert w[2, int(t0 * 1000) - 1, 44] == 0 and w[2, int((t0 + 0.03) * 1000) + 1, 44] == 1
    assert w[2, int(round((t0 + 0.015) * 1000)), 44] == pytest.approx(0.5, abs=1e-12)
    assert sim.mute_weights(np.array([2000.0])).shape == (1, 5, 1000, 70)


def test_imaging_memory_bound():
    assert fd.imaging_memory_bytes(24) < 24 * 5 * 1000 * 150 * 150 * 4 / 5


# -------------------------------------------------------------- invariants
def test_linear_in_wavelet_and_mirror_symmetric():
    sim = fd.OracleSim(pad=12, dtype=torch.float64, geom=SMALL)
    v = np.tile(np.linspace(1600, 3000, SMALL.nz)[:, None], (1, SMALL.nx))[None]
    a = sim.gathers(v)
    sim2 = fd.OracleSim(wavelet=2 * fd.ricker(SMALL), pad=12, dtype=torch.float64, geom=SMALL)
    np.testing.assert_array_equal(sim2.gathers(v), 2 * a)
    # shot 0 at column 0 and shot 2 at column nx-1 mirror each other for laterally constant v
    np.testing.assert_allclose(a[0, 2], a[0, 0, :, ::-1], rtol=0, atol=1e-12 * np.abs(a).max())


def test_reciprocity_homogeneous():
    geom = fd.Geometry(nz=24, nx=28, nt=260, ns=2, src=(5, 20))
    sim = fd.OracleSim(pad=12, dtype=torch.float64, geom=geom)
    g = sim.gathers(np.full((1, 24, 28), 2200.0))
    # source at 5 recorded at 20 equals source at 20 recorded at 5 (sponge border mask aside)
    assert rel(g[0, 0, :, 20], g[0, 1, :, 5]) < 1e-9


def _loss_grad_setup():
    sim = fd.OracleSim(pad=12, dtype=torch.float64, geom=SMALL, checkpoint_interval=40)
    v = np.full((1, SMALL.nz, SMALL.nx), 2000.0)
    truth = v.copy()
    truth[:, 8:] = 2400.0
    return sim, v, sim.gathers(truth)


def test_gradient_matches_finite_differences():
    sim, v, obs = _loss_grad_setup()
    grad = sim.gradient(v, obs)[0].sum(1)
    for loc in [(slice(10, 18), 14), (slice(5, 12), 0), (SMALL.nz - 1, slice(8, 16)), (SMALL.row, SMALL.src[1])]:
        d = np.zeros_like(v)
        d[(0,) + loc] = 1

        def loss(m):
            r = sim.gathers(m) - obs
            return 0.5 * np.sum(r * r)
        eps = 0.1
        num = (loss(v + eps * d) - loss(v - eps * d)) / (2 * eps)
        assert abs(num - np.sum(grad * d)) / max(abs(num), 1.0) < 1e-4


def test_checkpoint_interval_invariance():
    sim, v, obs = _loss_grad_setup()
    mute = np.array([2000.0])
    g1, i1 = sim.gradient(v, obs, v_mute=mute)
    other = fd.OracleSim(pad=12, dtype=torch.float64, geom=SMALL, checkpoint_interval=73)
    g2, i2 = other.gradient(v, obs, v_mute=mute)
    np.testing.assert_allclose(g2, g1, rtol=1e-10, atol=1e-8 * np.abs(g1).max())
    np.testing.assert_allclose(i2, i1, rtol=1e-12)


def test_dot_product_born_adjoint():
    """<y, F x> = <F* y, x> for the tangent-linear forward F and the adjoint gradient."""
    sim = fd.OracleSim(pad=12, dtype=torch.float64, geom=SMALL)
    rng = np.random.default_rng(5)
    v = 2000 + 300 * rng.random((2, SMALL.nz, SMALL.nx))
    x = rng.standard_normal(v.shape)
    y = rng.standard_normal((2, SMALL.ns, SMALL.nt, SMALL.nx))
    Fx = sim.born(v, x)
    _, ck, c2 = sim.forward(v, store=True)
    grad, _ = sim.adjoint(v, torch.tensor(y), ck, c2)
    lhs, rhs = float(np.sum(y * Fx)), float(np.sum(grad.numpy().sum(1) * x))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_born_matches_finite_differences():
    sim = fd.OracleSim(pad=12, dtype=torch.float64, geom=SMALL)
    rng = np.random.default_rng(6)
    v = 2000 + 300 * rng.random((1, SMALL.nz, SMALL.nx))
    dv = rng.standard_normal(v.shape)
    eps = 1e-3
    num = (sim.gathers(v + eps * dv) - sim.gathers(v - eps * dv)) / (2 * eps)
    assert rel(sim.born(v, dv), num) < 1e-7


# ------------------------------------------------------- CNN oracle pins
def _brute_conv3x3(x, w, b):
    """Same-padded 3x3 cross-correlation by explicit loops (tiny sizes)."""
    N, C, H, W = x.shape
    out = np.zeros((N, w.shape[0], H, W))
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    for o in range(w.shape[0]):
        acc = np.full((N, H, W), b[o])
        for c in range(C):
            for dy in range(3):
                for dx in range(3):
                    acc += w[o, c, dy, dx] * xp[:, c, dy:dy + H, dx:dx + W]
        out[:, o] = acc
    return out


def test_cnn_oracle_forward_is_cross_correlation():
    from oracle import cnn as ocnn
    layers = ((2, 3, True), (3, 1, False))
    rng = np.random.default_rng(0)
    n = sum(ci * co * 9 + co for ci, co, _ in layers)
    flat = torch.tensor(rng.standard_normal(n))
    x = rng.standard_normal((2, 2, 5, 6))
    (w1, b1, _), (w2, b2, _) = ocnn.unflatten(flat, layers)
    h = np.maximum(_brute_conv3x3(x, w1.numpy(), b1.numpy()), 0)
    want = _brute_conv3x3(h, w2.numpy(), b2.numpy())
    got = ocnn.forward(flat, torch.tensor(x), layers).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # a flipped (true convolution) kernel would differ
    flipped = _brute_conv3x3(x, w1.numpy()[:, :, ::-1, ::-1], b1.numpy())
    assert np.abs(flipped - _brute_conv3x3(x, w1.numpy(), b1.numpy())).max() > 1e-3


def test_cnn_oracle_gradient_finite_differences():
    from oracle import cnn as ocnn
    layers = ((2, 4, True), (4, 4, True), (4, 1, False))
    n = sum(ci * co * 9 + co for ci, co, _ in layers)
    g = torch.Generator().manual_seed(1)
    flat = torch.randn(n, generator=g, dtype=torch.float64) * 0.5
    x = torch.randn(2, 2, 7, 6, generator=g, dtype=torch.float64)
    t = torch.randn(2, 1, 7, 6, generator=g, dtype=torch.float64)
    _, grad = ocnn.loss_and_grad(flat, x, t, layers)
    idx = torch.randperm(n, generator=g)[:25]
    h = 1e-6
    for i in idx.tolist():
        e = torch.zeros(n, dtype=torch.float64)
        e[i] = h
        lp, _ = ocnn.loss_and_grad(flat + e, x, t, layers)
        lm, _ = ocnn.loss_and_grad(flat - e, x, t, layers)
        fd_ = (lp - lm) / (2 * h)
        assert abs(fd_ - float(grad[i])) <= 1e-6 * max(1.0, abs(fd_)), (i, fd_, float(grad[i]))


def test_cnn_oracle_adam_closed_form():
    from oracle import cnn as ocnn
    p = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    g1 = torch.tensor([0.1, -0.3, 0.0], dtype=torch.float64)
    m = v = torch.zeros(3, dtype=torch.float64)
    p1, m1, v1 = ocnn.adam(p, g1, m, v, 1, lr=0.01)
    # step 1: bias-corrected m/sqrt(v) = sign(g) (for g != 0)
    np.testing.assert_allclose(p1.numpy(), [1.0 - 0.01 * 0.1 / (0.1 + 1e-8), -2.0 + 0.01 * 0.3 / (0.3 + 1e-8), 0.5],
                               rtol=0, atol=1e-15)
    g2 = torch.tensor([0.2, 0.1, -0.4], dtype=torch.float64)
    p2, _, _ = ocnn.adam(p1, g2, m1, v1, 2, lr=0.01)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(p2.numpy(), (p1 - 0.01 * mh / (vh.sqrt() + 1e-8)).numpy(), rtol=0, atol=1e-14)


# ---------------------------------------------------- OpenFWI score oracle pins
def _brute_ssim(a, b, size=11, sigma=1.5):
    """SSIM from the published formula with explicit loops and zero padding."""
    H, W = a.shape
    r = size // 2
    g = np.array([np.exp(-(k - r) ** 2 / (2 * sigma ** 2)) for k in range(size)])
    g /= g.sum()
    tot = 0.0
    for y in range(H):
        for x in range(W):
            m = {'a': 0.0, 'b': 0.0, 'aa': 0.0, 'bb': 0.0, 'ab': 0.0}
            for dy in range(size):
                for dx in range(size):
                    yy, xx = y + dy - r, x + dx - r
                    if 0 <= yy < H and 0 <= xx < W:
                        wgt = g[dy] * g[dx]
                        va, vb = a[yy, xx], b[yy, xx]
                        m['a'] += wgt * va
                        m['b'] += wgt * vb
                        m['aa'] += wgt * va * va
                        m['bb'] += wgt * vb * vb
                        m['ab'] += wgt * va * vb
            sa, sb, sab = m['aa'] - m['a'] ** 2, m['bb'] - m['b'] ** 2, m['ab'] - m['a'] * m['b']
            c1, c2 = 1e-4, 9e-4
            tot += ((2 * m['a'] * m['b'] + c1) * (2 * sab + c2)) / ((m['a'] ** 2 + m['b'] ** 2 + c1) * (sa + sb + c2))
    return tot / (H * W)


def test_score_oracle_ssim_matches_brute_force():
    from oracle import scores
    rng = np.random.default_rng(3)
    a = rng.uniform(0, 1, (9, 13))
    b = np.clip(a + 0.1 * rng.standard_normal((9, 13)), 0, 1)
    assert abs(scores.ssim(a, b) - _brute_ssim(a, b)) < 1e-13
    assert abs(scores.ssim(a, a) - 1.0) < 1e-14
    # a planted wrong sigma is caught
    assert abs(_brute_ssim(a, b, sigma=2.0) - scores.ssim(a, b)) > 1e-4


def test_score_oracle_metrics_closed_forms():
    from oracle import scores
    rng = np.random.default_rng(4)
    v = rng.uniform(1500, 4500, (12, 10))
    m = scores.metrics(v, v)
    assert m['mae'] == 0 and m['rmse'] == 0 and abs(m['ssim'] - 1) < 1e-14
    m = scores.metrics(v + 300.0, v)        # +300 m/s = +0.2 in the [-1, 1] scale
    assert abs(m['mae'] - 0.2) < 1e-12 and abs(m['rmse'] - 0.2) < 1e-12
    np.testing.assert_allclose(scores.norm([1500.0, 3000.0, 4500.0]), [-1.0, 0.0, 1.0])
