"""Benchmark: batched forward + exact-adjoint RTM images on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--models 64] [--seed 0] [--no-cpu]
    python bench.py --workload mixed [--models-per-size 32]    (uneven batch, see run_mixed)
    python bench.py --workload autograd                        (differentiation, see run_autograd)

Workload (one "step"): 64 seeded 70 x 70 models of mixed structure (curved
layers, a planar fault, smoothed texture) x 3 backgrounds (surface velocity +
0.6/1.0/1.4 x depth trend) x 5 shots = 960 shot-images; nt = 1000, dt = 1 ms,
dx = 10 m, pad 40 (150 x 150 padded grid), 15 Hz Ricker, FP32.  One step runs
the forward with the Laplacian history, the muted residual, the exact adjoint
with the imaging condition, the gradient fold and the image epilogue, on
inputs resident in HBM (observed gathers, coefficient planes, mute velocities).
Model generation, observed-data generation and host coefficient setup are
outside the timed region.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(HIP events on the launching stream) and the host-CPU baseline (the eager
PyTorch oracle, timed on a bounded sample in this process).
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

METRIC = 'RTM shot-images/s, 70x70 nt=1000 fwd+adjoint; learned vs plain refine misfit'
PEAK_HBM_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 16          # read p(t-1), p(t), c2; write p(t+1); FP32


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel of the batched workload, from the two
    rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB copy,
    recorded in profiles/hbm_traffic.json; None when not recorded."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'hbm_traffic.json')
    try:
        with open(path) as f:
            return json.load(f)['kernels'][kernel]['bytes']
    except (OSError, KeyError, ValueError):
        return None


def measured_lds(kernel, kernel_ms):
    """LDS bytes per launch of a kernel of the batched workload (rocprofv3
    SQ_INSTS_LDS_LOAD_BANDWIDTH + SQ_INSTS_LDS_STORE_BANDWIDTH, 64-byte units)
    and its issue figures, from profiles/lds_valu.json, as an LDS roofline
    beside the HBM one; None when not recorded."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'lds_valu.json')
    try:
        with open(path) as f:
            rec = json.load(f)
        k = rec['kernels'][kernel]
        achieved = k['lds_bytes'] / (kernel_ms * 1e-3) / 1e9
        return {'bytes': k['lds_bytes'], 'achieved': achieved, 'peak': rec['peak_lds_GBps'], 'unit': 'GB/s',
                'frac': achieved / rec['peak_lds_GBps'], 'valu_busy': k['valu_busy'], 'lds_util': k['lds_util'],
                'source': 'profiles/lds_valu.json (rocprofv3 LDS bandwidth and issue counters)'}
    except (OSError, KeyError, ValueError):
        return None


def workload(n_models, seed, sim):
    """Models, observed gathers (device), backgrounds and mute velocities."""
    from shotsweep_amd import background as BG, synth
    models = synth.mixed(n_models, seed=seed)
    out, _, _ = sim.forward(models)
    obs = out.contiguous()
    obs_host = obs.cpu().numpy()
    trend = BG.depth_trend(models)
    v_s, fallbacks = [], 0
    for k, est in enumerate(BG.estimate_surface_velocity_device(obs)):   # picks on the GPU, fits on the host
        if isinstance(est, ValueError):
            v_s.append(float(models[k, 0].mean()))
            fallbacks += 1
        else:
            v_s.append(est['velocity'])
    v_s = np.asarray(v_s)
    bgs = np.concatenate([BG.backgrounds(v, trend) for v in v_s])
    v_mute = np.repeat(v_s, len(BG.RHOS))
    return models, obs, obs_host, bgs, v_mute, fallbacks


def cpu_baseline(bgs, obs_host, v_mute, n_rows):
    """Time the oracle (eager PyTorch CPU restatement of the reference) on n_rows backgrounds.

    One untimed warm-up on 3 rows, then three timed runs on consecutive thirds
    of the sample; the rate is the median of the three.  Also returns the FP64
    oracle's images of the first model (3 rows)."""
    from oracle import fd
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    osim = fd.OracleSim()
    obs_rep = np.repeat(obs_host, 3, 0)[:n_rows]
    osim.images(bgs[:3], obs_rep[:3], v_mute=v_mute[:3])
    part = max(1, n_rows // 3)
    imgs, illums, rates, times = [], [], [], []
    for k in range(3):
        sl = slice(k * part, n_rows if k == 2 else (k + 1) * part)
        t0 = time.perf_counter()
        img, illum = osim.images(bgs[sl], obs_rep[sl], v_mute=v_mute[sl])
        dt = time.perf_counter() - t0
        imgs.append(img)
        illums.append(illum)
        rates.append((sl.stop - sl.start) * 5 / dt)
        times.append(dt)
    shot_images = n_rows * 5
    o64 = fd.OracleSim(dtype=torch.float64)
    img64, _ = o64.images(bgs[:3], obs_rep[:3].astype(np.float64), v_mute=v_mute[:3])
    return {'value': float(np.median(rates)), 'unit': 'shot-images/s', 'cores': cores, 'kind': 'port',
            'sample': f'{n_rows} background rows ({shot_images} shot-images) of the same workload in three timed '
                      f'runs after a warm-up, oracle FP32 eager PyTorch on {cores} threads, median of '
                      f'{", ".join(f"{t:.1f}" for t in times)} s runs'}, \
        np.concatenate(imgs), np.concatenate(illums), img64


def mixed_workload(n_per_size, seed, acqs):
    """Per grid size: models, observed gathers (device), backgrounds and mute velocities.

    Surface velocity = mean of the model's top row (the first-break picker is
    specific to the 70-receiver acquisition); trend = depth_trend of the size's
    models; 3 backgrounds per model as in the batched workload."""
    from shotsweep_amd import background as BG, gpu, synth
    out = []
    for k, acq in enumerate(acqs):
        sim = gpu.Sim(geom=acq)
        models = synth.mixed(n_per_size, seed=seed + 1000 * k, nz=acq.nz, nx=acq.nx)
        obs = torch.cat([sim.forward(models[i:i + 8])[0] for i in range(0, n_per_size, 8)]).contiguous()
        sim.release()
        trend = BG.depth_trend(models)
        v_s = models[:, 0].mean(1)
        bgs = np.concatenate([BG.backgrounds(v, trend, nx=acq.nx) for v in v_s])
        vm = np.repeat(v_s, len(BG.RHOS))
        out.append({'v_bg': bgs, 'obs': obs, 'obs_repeat': len(BG.RHOS),
                    'v_mute': torch.tensor(vm, dtype=torch.float32, device='cuda')})
    return out


def mixed_cpu_baseline(acqs, batches, shot_images):
    """Oracle on one background row per grid size, extrapolated to the whole workload."""
    from oracle import fd
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    total, parts, imgs = 0.0, [], []
    for acq, b in zip(acqs, batches):
        osim = fd.OracleSim(geom=fd.Geometry(nz=acq.nz, nx=acq.nx, nt=acq.nt, ns=acq.ns, row=acq.row))
        obs = b['obs'][:1].cpu().numpy()
        t0 = time.perf_counter()
        img, illum = osim.images(b['v_bg'][:1], obs, v_mute=b['v_mute'][:1].cpu().numpy())
        dt = time.perf_counter() - t0
        total += dt * len(b['v_bg'])
        parts.append(f'{acq.nz}x{acq.nx}: {dt:.1f} s')
        imgs.append((img, illum))
    return {'value': shot_images / total, 'unit': 'shot-images/s', 'cores': cores, 'kind': 'port',
            'sample': 'one background row per grid size (oracle FP32 eager PyTorch on '
                      f'{cores} threads; {", ".join(parts)}), extrapolated linearly to every row'}, imgs


def run_mixed(args):
    """Uneven batch: 32 models each of 70x70 (nt 1000, 5 shots), 100x140 (nt 1500,
    7 shots) and 200x200 (nt 2000, 10 shots) x 3 backgrounds in one mixed sweep."""
    from shotsweep_amd.mixed import MixedSweep, config_acquisitions
    acqs = config_acquisitions()
    batches = mixed_workload(args.models_per_size, args.seed, acqs)
    ms = MixedSweep(acqs)
    shot_images = sum(len(b['v_bg']) * a.ns for a, b in zip(acqs, batches))
    rows = [len(b['v_bg']) for b in batches]
    nchunks, _ = ms._chunks(rows, [b['obs_repeat'] for b in batches])
    print(f'[bench] mixed: rows {rows}, {shot_images} shot-images, {nchunks} chunks', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        ms.images_device(batches)
    torch.cuda.synchronize()
    print(f'[bench] warmup (with autotuning) {time.perf_counter() - t0:.1f} s, tiles {ms.tiles}',
          file=sys.stderr, flush=True)
    ms.events = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ms.images_device(batches)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    events, ms.events = ms.events, None
    ms_per_step = wall / args.steps * 1e3
    value = shot_images / (wall / args.steps)
    per = {}
    for name, a, b in events:
        per[name] = per.get(name, 0.0) + a.elapsed_time(b) / args.steps
    dominant = max(per, key=per.get)
    cell_updates = sum(len(b['v_bg']) * a.ns * a.nt * (a.nz + 80) * (a.nx + 80) for a, b in zip(acqs, batches))
    alg_bytes = BYTES_PER_CELL_UPDATE * cell_updates
    achieved = alg_bytes / (per[dominant] * 1e-3) / 1e9
    print(f'[bench] {ms_per_step:.1f} ms/step, sweep ms per step {per}', file=sys.stderr, flush=True)
    result = {
        'metric': METRIC, 'value': value, 'unit': 'shot-images/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded mixed-structure models at each size)',
        'config': {'workload': 'uneven batched RTM: 32 models each of 70x70 (nt 1000, 5 shots), 100x140 '
                               '(nt 1500, 7 shots), 200x200 (nt 2000, 10 shots) x 3 backgrounds, pad 40, '
                               'dx 10 m, 15 Hz Ricker', 'models_per_size': args.models_per_size,
                   'shot_images': shot_images, 'seed': args.seed, 'chunks': nchunks,
                   'tiles': {f'{a.nz}x{a.nx}': {('forward', 'adjoint')[d]: t for d, t in ms.tiles[k].items()}
                             for k, a in enumerate(acqs)}},
        'roofline': {'bound': 'hbm', 'kernel': dominant, 'achieved': achieved, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                     'frac': achieved / PEAK_HBM_GBS, 'traffic': None, 'sweep_ms_per_step': per,
                     'algorithmic_bytes_per_direction': alg_bytes},
    }
    if not args.no_cpu:
        base, imgs = mixed_cpu_baseline(acqs, batches, shot_images)
        result['cpu_baseline'] = base
        result['speedup_vs_cpu'] = value / base['value']
        errs = []
        for (cimg, cill), (gimg, gill) in zip(imgs, out):
            g = gimg[:1].cpu().numpy()
            errs.append(float(np.linalg.norm(g - cimg) / np.linalg.norm(cimg)))
        result['image_rel_l2_vs_cpu_oracle'] = max(errs)
    print(json.dumps(result), flush=True)


def run_autograd(args):
    """Differentiation workload: misfit 0.5 * sum((F(v) - obs)^2) and its gradient through
    the autograd Function for 16 models x 5 shots (70 x 70, nt 1000), start = smoothed truth
    (sigma 8), checkpoints every 50 steps with recompute (the reference's scheme)."""
    from shotsweep_amd import autograd, gpu, synth
    sim = gpu.Sim(device='cuda', checkpoint_interval=50, history='checkpoint')
    true = synth.mixed(16, seed=args.seed)
    obs = sim.forward(true)[0].contiguous()
    start = synth.smooth(true, 8.0)
    ns = sim.geom.ns
    for _ in range(args.warmup):
        autograd.misfit_and_gradient(start, obs, sim)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, grad = autograd.misfit_and_gradient(start, obs, sim)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    # kernel breakdown from one more step with HIP events around every launch (the
    # segment recomputes then run on the caller stream, so each interval is one launch)
    sim.timing, sim.timing_log = True, []
    autograd.misfit_and_gradient(start, obs, sim)
    sim.timing = False
    per = {'forward': sum(t['forward'][1] for t in sim.timing_log),
           'adjoint': sum(t['adjoint'][1] for t in sim.timing_log)}
    dominant = max(per, key=per.get)
    units = len(true) * ns
    alg = BYTES_PER_CELL_UPDATE * sim.H * sim.W * sim.geom.nt * units
    result = {
        'metric': METRIC, 'value': units / wall, 'unit': 'shot-gradients/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': wall * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded mixed-structure models)',
        'config': {'workload': 'differentiation: misfit gradient of 16 mixed 70x70 models x 5 shots, nt 1000, '
                               'autograd.Function, checkpoints every 50 steps with recompute',
                   'models': len(true), 'seed': args.seed, 'misfit': loss},
        'roofline': {'bound': 'hbm', 'kernel': dominant, 'achieved': alg / (per[dominant] * 1e-3) / 1e9,
                     'peak': PEAK_HBM_GBS, 'unit': 'GB/s', 'frac': alg / (per[dominant] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     'traffic': None, 'kernel_ms_per_step': per, 'algorithmic_bytes_per_direction': alg,
                     'note': 'forward time includes the per-segment recompute (two forward passes); kernel '
                             'times from one extra step with events around every launch, in which the '
                             'segment recomputes run on the caller stream (the timed steps run each recompute '
                             'on a second stream beside the previous segment\'s adjoint)'},
    }
    if not args.no_cpu:
        from oracle import fd
        cores = len(os.sched_getaffinity(0))
        torch.set_num_threads(cores)
        osim = fd.OracleSim(checkpoint_interval=50)
        t0 = time.perf_counter()
        og, _ = osim.gradient(start[:1], obs[:1].cpu().numpy())
        dt = time.perf_counter() - t0
        result['cpu_baseline'] = {'value': ns / dt, 'unit': 'shot-gradients/s', 'cores': cores, 'kind': 'port',
                                  'sample': f'one model (5 shot gradients), oracle FP32 eager PyTorch with '
                                            f'checkpoints every 50 steps on {cores} threads, {dt:.1f} s'}
        result['speedup_vs_cpu'] = result['value'] / result['cpu_baseline']['value']
        g0 = grad[0].cpu().numpy().astype(np.float64)
        want = og.sum(1)[0].astype(np.float64)
        result['gradient_rel_l2_vs_cpu_oracle'] = float(np.linalg.norm(g0 - want) / np.linalg.norm(want))
    print(json.dumps(result), flush=True)


def copy_probe(nbytes=1 << 30, reps=3):
    """Plain device copies of known size (read nbytes, write nbytes each): the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    x = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({'copy_probe_bytes': nbytes, 'reps': reps}), flush=True)


def bandwidth_probe(nbytes=8 << 30, reps=5):
    """Device fill (write only), copy (read + write) and sum (read only) rates of
    an nbytes float32 buffer, timed with events: reference points for the
    history stream of the FD kernels."""
    x = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    out = {}
    for name, fn, moved in (('fill', lambda: x.fill_(1.0), nbytes), ('copy', lambda: y.copy_(x), 2 * nbytes),
                            ('sum', lambda: x.sum(), nbytes)):
        fn()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps):
            fn()
        t1.record()
        torch.cuda.synchronize()
        out[name + '_TBps'] = moved * reps / (t0.elapsed_time(t1) * 1e-3) / 1e12
    print(json.dumps({'bandwidth_probe_bytes': nbytes, **out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--models', type=int, default=64)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--history', default='auto', choices=('auto', 'full', 'checkpoint'))
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-models', type=int, default=20,
                    help='models in the bounded host-CPU baseline sample (x3 backgrounds x5 shots)')
    ap.add_argument('--workload', default='batched', choices=('batched', 'mixed', 'autograd'),
                    help='batched: 64 models of 70x70 (the headline metric); mixed: the uneven batch; '
                         'autograd: the differentiation workload')
    ap.add_argument('--models-per-size', type=int, default=32)
    ap.add_argument('--copy-probe', action='store_true', help='only run 1 GiB device copies (counter scaling)')
    ap.add_argument('--bandwidth-probe', action='store_true', help='only time 8 GiB device fill / copy / sum')
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.bandwidth_probe:
        return bandwidth_probe()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: the path shards by model with no exchange in the time loop; '
              'multi-GPU runs are not part of this build', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'mixed':
        return run_mixed(args)
    if args.workload == 'autograd':
        return run_autograd(args)
    from shotsweep_amd import gpu

    dev = torch.device('cuda')
    sim = gpu.Sim(device=dev, history=args.history)
    models, obs, obs_host, bgs, v_mute, fallbacks = workload(args.models, args.seed, sim)
    prep = sim._prepare(bgs)
    vm = sim._tensor(v_mute).contiguous()
    nb = prep['nb']
    shot_images = nb * sim.geom.ns
    print(f'[bench] {args.models} models -> {nb} backgrounds -> {shot_images} shot-images; '
          f'history segment {sim._segment(nb)}; surface-velocity fallbacks {fallbacks}', file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        sim.images_device(prep, obs, 3, vm)
    torch.cuda.synchronize()
    sim.timing, sim.timing_log = True, []
    per_kernel = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        img, illum = sim.images_device(prep, obs, 3, vm)
        for name, (n, ms) in sim.timing_log[-1].items():
            per_kernel.setdefault(name, []).append(ms / max(n, 1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sim.timing = False
    ms_per_step = wall / args.steps * 1e3
    value = shot_images / (wall / args.steps)

    avg = {k: float(np.mean(v)) for k, v in per_kernel.items()}
    dominant = max(avg, key=avg.get)
    H, W, nt = sim.H, sim.W, sim.geom.nt
    alg_bytes = BYTES_PER_CELL_UPDATE * H * W * nt * shot_images       # one cell-update per cell-step per launch
    achieved = alg_bytes / (avg[dominant] * 1e-3) / 1e9
    print(f'[bench] {ms_per_step:.1f} ms/step, kernels {avg}', file=sys.stderr, flush=True)

    result = {
        'metric': METRIC, 'value': value, 'unit': 'shot-images/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded mixed-structure models)',
        'config': {'workload': 'batched RTM: 64 mixed 70x70 models x 3 backgrounds x 5 shots = 960 shot-images, '
                               'nt=1000, dt=1 ms, dx=10 m, pad 40 (150x150), 15 Hz Ricker',
                   'models': args.models, 'shot_images': shot_images, 'seed': args.seed,
                   'history_segment': sim._segment(nb)},
        'roofline': {'bound': 'hbm', 'kernel': dominant, 'achieved': achieved, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                     'frac': achieved / PEAK_HBM_GBS, 'traffic': measured_traffic(dominant),
                     'traffic_source': 'profiles/hbm_traffic.json (rocprofv3 FETCH_SIZE x2.0, WRITE_SIZE x1.0)',
                     'kernel_ms': avg, 'algorithmic_bytes_per_launch': alg_bytes,
                     'lds': measured_lds(dominant, avg[dominant])},
    }
    # first-break picking for the surface-velocity anchors: HIP kernel vs the host NumPy picks
    from shotsweep_amd import background as BG
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    BG.estimate_surface_velocity_device(obs)
    dev_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(len(obs_host)):
        try:
            BG.estimate_surface_velocity(obs_host[k].astype(np.float64))
        except ValueError:
            pass
    result['surface_picking'] = {'models': len(obs_host), 'device_s': dev_s,
                                 'host_s': time.perf_counter() - t0,
                                 'note': 'device: ss_first_breaks picks + host fits; host: NumPy picks + fits'}
    # the drop-in call with host buffers (NumPy in, NumPy out): PCIe-inclusive, never `value`
    obs_rep = np.repeat(obs_host, 3, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.images(bgs, obs_rep, v_mute=v_mute)
    host_s = time.perf_counter() - t0
    result['host_buffers'] = {'call': 'Sim.images(v_bg, obs, v_mute) with NumPy inputs and outputs',
                              'seconds': host_s, 'shot_images_per_s': shot_images / host_s}
    if not args.no_cpu:
        rows = 3 * args.cpu_models             # cpu_models models x 3 backgrounds x 5 shots
        base, cimg, cillum, cimg64 = cpu_baseline(bgs, obs_host, v_mute, rows)
        gimg = img[:rows].cpu().numpy()
        err = float(np.linalg.norm(gimg - cimg) / np.linalg.norm(cimg))
        result['cpu_baseline'] = base
        result['speedup_vs_cpu'] = value / base['value']
        result['image_rel_l2_vs_cpu_oracle'] = err
        result['illum_bitwise_equal_cpu_oracle'] = bool(np.array_equal(illum[:rows].cpu().numpy(), cillum))
        g3 = gimg[:3].astype(np.float64)
        result['image_rel_l2_vs_fp64_oracle'] = float(np.linalg.norm(g3 - cimg64) / np.linalg.norm(cimg64))
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
