"""bench.py's command line on the CPU: a multi-GPU request exits with a message
and touches no device."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    env = dict(os.environ, WORLD_SIZE='1')
    r = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--gpus', '2', '--steps', '1'],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and 'measures one GPU' in r.stderr and r.stdout == ''


def test_cpu_baseline_is_median_of_three_timed_runs():
    """bench.cpu_baseline (the oracle on the host) on one model: a warm-up, three
    timed runs over thirds of the sample, the median rate, the sample's images
    in row order, and the FP64 oracle's images of the first model."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    import bench
    from oracle import fd
    from shotsweep_amd import background as BG, synth
    models = synth.mixed(1, seed=3)
    obs = fd.OracleSim().gathers(models)
    trend = BG.depth_trend(models)
    bgs = BG.backgrounds(float(models[0, 0].mean()), trend)
    vm = np.repeat(models[:, 0].mean(1), 3)
    base, img, illum, img64 = bench.cpu_baseline(bgs, obs, vm, 3)
    assert base['kind'] == 'port' and base['unit'] == 'shot-images/s' and base['cores'] >= 1
    assert base['value'] > 0 and 'median' in base['sample'] and '3 background rows' in base['sample']
    assert img.shape == illum.shape == (3, 5, 70, 70) and img64.shape == (3, 5, 70, 70)
    want, _ = fd.OracleSim().images(bgs, np.repeat(obs, 3, 0), v_mute=vm)
    np.testing.assert_array_equal(img, want)       # thirds of the sample, concatenated in order
    e = np.linalg.norm(img.astype(np.float64) - img64) / np.linalg.norm(img64)
    assert 0 < e < 1e-4                             # FP32 against FP64 rounding, not a different result


def test_board_power_sampler(tmp_path):
    """bench.BoardPower reads a hwmon power file (microwatts) on a host thread while
    open and reports mean and maximum watts after the first quarter of samples, the
    power cap beside it; without a readable file it reports nothing."""
    import time
    sys.path.insert(0, str(ROOT))
    import bench
    f = tmp_path / 'power1_input'
    f.write_text('1250000000\n')
    (tmp_path / 'power1_cap').write_text('1400000000\n')
    p = bench.BoardPower()
    p.path = str(f)
    with p:
        time.sleep(0.15)
    s = p.summary()
    assert s['mean_W'] == s['max_W'] == 1250.0 and s['cap_W'] == 1400.0 and s['samples'] >= 3
    q = bench.BoardPower()
    q.path = str(tmp_path / 'missing')
    with q:
        time.sleep(0.05)
    assert q.summary() is None
