"""Image sweep of the processing harness (2_process.py:193-214)."""
import numpy as np
import pytest
import torch

from oracle import fd
from shotsweep_amd import background as BG
from shotsweep_amd import synth
from shotsweep_amd.sweep import postprocess


def test_postprocess_closed_form():
    img = np.full((2, 3, 5, 70, 70), 2.0, np.float32)
    img[1] *= -3
    img[1, 0, 0, 10, 10] = 1e6          # one outlier above the 99th percentile is clipped to -4..4
    out = postprocess(img)
    assert out.dtype == np.float16 and out.shape == img.shape
    assert np.all(out[..., :2, :] == 0)
    assert np.all(out[0, ..., 2:, :] == 1) and out[1, 0, 0, 10, 10] == 4
    assert np.all(out[1, 1, ..., 2:, :] == -1)


def test_postprocess_rejects_zero_and_nonfinite():
    with pytest.raises(ValueError, match='zero-energy'):
        postprocess(np.zeros((1, 3, 5, 70, 70)))
    bad = np.ones((1, 3, 5, 70, 70))
    bad[0, 0, 0, 5, 5] = np.nan
    with pytest.raises(ValueError, match='nonfinite'):
        postprocess(bad)


@pytest.mark.gpu
def test_sweep_matches_oracle_pipeline():
    from shotsweep_amd import gpu
    from shotsweep_amd.sweep import sweep_batch
    print('\n[case] sweep of 2 models vs the oracle pipeline', flush=True)
    torch.set_num_threads(8)
    models = np.stack([synth.flat_layers(seed=3), synth.mixed(1, seed=5)[0]])
    osim = fd.OracleSim()
    obs = osim.gathers(models).astype(np.float64)
    trend = BG.depth_trend(synth.mixed(8, seed=1))
    out = sweep_batch(gpu.Sim(device='cuda'), obs, trend)
    # the oracle pipeline, following 2_process.py:196-210
    v1 = np.array([BG.estimate_surface_velocity(o)['velocity'] for o in obs])
    np.testing.assert_array_equal(out['surface_velocity'], v1.astype(np.float32))
    bgs = np.concatenate([BG.backgrounds(v, trend) for v in v1])
    img, _ = osim.images(bgs, np.repeat(obs, 3, 0), v_mute=np.repeat(v1, 3))
    want = postprocess(img.reshape(2, 3, 5, 70, 70))
    e = np.linalg.norm(out['raw'].ravel() - img.ravel()) / np.linalg.norm(img)
    diff = np.abs(out['images'].astype(np.float32) - want.astype(np.float32))
    print(f'  raw image rel L2 {e:.3e}; fp16 entries differing: {(diff > 0).mean():.2e}, max {diff.max():.2e}')
    assert e < 1e-5
    # rounding to float16 moves each value by at most half an ulp, so two cache entries
    # differ by at most one ulp (of the larger) plus the scaled raw difference; near the
    # zeroed surface rows the Laplacian filter cancels, so a few tiny pixels carry a
    # larger relative raw difference than the 1e-5 image-wide bound suggests
    raw = out['raw'].reshape(want.shape).astype(np.float64)
    ref = img.reshape(want.shape).astype(np.float64)
    raw[..., :2, :] = 0
    ref[..., :2, :] = 0
    scale = np.percentile(np.abs(ref).reshape(2, -1), 99, axis=1)[:, None, None, None, None]
    big = np.maximum(np.abs(want.astype(np.float32)), np.abs(out['images'].astype(np.float32)))
    ulp = np.spacing(np.maximum(big, np.float32(2.0 ** -14)).astype(np.float16)).astype(np.float64)
    bound = ulp + np.abs(raw - ref) / scale * (1 + 1e-3) + big * 1e-5   # last term: the two p99 scales
    assert np.all(diff <= bound) and (diff > 0).mean() < 0.02
