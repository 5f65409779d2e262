"""Benchmark: batched forward + exact-adjoint RTM images on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--models 64] [--seed 0] [--no-cpu] [--legs all|none|...]
    python bench.py --workload mixed|autograd|refine   (one leg alone, its own JSON line)

Headline workload (one "step"): 64 seeded 70 x 70 models of mixed structure
(curved layers, a planar fault, smoothed texture) x 3 backgrounds (surface
velocity + 0.6/1.0/1.4 x depth trend) x 5 shots = 960 shot-images; nt = 1000,
dt = 1 ms, dx = 10 m, pad 40 (150 x 150 padded grid), 15 Hz Ricker, FP32.  One
step is one native call (ss_rtm_f32): the forward with the Laplacian history,
the muted residual, the exact adjoint with the imaging condition, the gradient
fold and the image epilogue, on inputs resident in HBM.  `value` is the median
over the timed steps (each bracketed by a synchronize).  The `protocol` object
times the same call with the observed gathers copied from pinned host memory
and the images and illumination copied back inside the timed region (median of
three runs).  Model generation, observed-data generation and the host
surface-velocity fits are outside every timed region.

The roofline of the dominant kernel: its average launch time comes from HIP
events recorded on the launching stream during the timed steps; its VALU
instruction count, HBM traffic and LDS figures come from the rocprofv3 counter
passes of the same workload (profiles/rtm_batched_counters.json).  The kernels
keep the wavefield on chip, so the binding ceiling is instruction issue (one
64-lane VALU instruction per clock per CU: 256 CUs x 2.4 GHz), reported as
`frac`; the measured HBM traffic and the 16-byte-per-cell-update algorithmic
rate sit beside it.

The `legs` object holds the other configurations, each with its own host-CPU
baseline (the oracle) and its error against the oracle: `differentiation`
(16 models, checkpoints every 50 steps, autograd), `uneven` (32 models at each
of 70x70 / 100x140 / 200x200 with their step and shot counts, one mixed sweep)
and `refine` (the learned preconditioner against the plain step on a bounded
sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = 'RTM shot-images/s, 70x70 nt=1000 fwd+adjoint; learned vs plain refine misfit'
PEAK_HBM_GBS = 8000.0
CUS, PEAK_CLOCK_GHZ = 256, 2.4
PEAK_VALU_TOPS = CUS * 64 * PEAK_CLOCK_GHZ / 1e3      # 64-lane VALU instructions: 39.3 T lane-ops/s
BYTES_PER_CELL_UPDATE = 16          # read p(t-1), p(t), c2; write p(t+1); FP32


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def counters():
    """Per-kernel counter figures of the batched workload (rocprofv3 passes), or None."""
    try:
        with open(os.path.join(ROOT, 'profiles', 'rtm_batched_counters.json')) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def roofline(kernel, kernel_ms, alg_bytes):
    """Roofline of one launch of `kernel` (HIP-event average kernel_ms) against its
    binding ceiling, VALU issue; HBM and LDS fractions and the algorithmic rate beside it."""
    t = kernel_ms * 1e-3
    out = {'bound': 'valu', 'kernel': kernel, 'achieved': None, 'peak': PEAK_VALU_TOPS, 'unit': 'TOP/s',
           'frac': None, 'traffic': None, 'kernel_ms': kernel_ms,
           'algorithmic_bytes_per_launch': alg_bytes, 'algorithmic_rate_TBps': alg_bytes / t / 1e12,
           'note': 'VALU lane-ops (64 per instruction) per second against 256 CUs x 64 lanes x 2.4 GHz; '
                   'traffic = measured HBM bytes per launch (FETCH_SIZE x2 + WRITE_SIZE, scaled by a 1 GiB '
                   'copy); the 16 B/cell-update algorithmic rate exceeds the HBM peak because the wavefield '
                   'stays in LDS and registers.  The launches run at the board power cap (board_power), '
                   'which holds the clock below 2.4 GHz: counter_run gives the clock and the VALU-busy '
                   'fraction at that clock'}
    rec = counters()
    k = rec and rec['kernels'].get(kernel)
    if not k:
        return out
    achieved = k['valu_instructions'] * 64 / t / 1e12
    out.update(achieved=achieved, frac=achieved / PEAK_VALU_TOPS, traffic=k['hbm_bytes'],
               valu_instructions_per_launch=k['valu_instructions'], valu_per_pair_step=k['valu_per_pair_step'],
               hbm={'achieved': k['hbm_bytes'] / t / 1e9, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                    'frac': k['hbm_bytes'] / t / 1e9 / PEAK_HBM_GBS},
               lds={'bytes': k['lds_bytes'], 'achieved': k['lds_bytes'] / t / 1e9, 'peak': rec['peak']['lds_GBps'],
                    'unit': 'GB/s', 'frac': k['lds_bytes'] / t / 1e9 / rec['peak']['lds_GBps'],
                    'bank_conflict_share': k['lds_bank_conflict_share']},
               counter_run={'clock_GHz': k['clock_GHz'], 'valu_busy': k['valu_busy'], 'ms': k['counter_run_ms']},
               source='profiles/rtm_batched_counters.json')
    return out


def sweep_roofline(acqs, rows, sweep_ms, alg_bytes):
    """Roofline of the uneven batch's adjoint sweep (its slower direction): VALU lane-ops
    per step from the per-pair-step instruction counts of each grid size's adjoint
    kernel (profiles/sweep_uneven_counters.json) over the sweep's HIP-event time."""
    t = sweep_ms * 1e-3
    out = {'bound': 'valu', 'kernel': 'sweep_adjoint', 'achieved': None, 'peak': PEAK_VALU_TOPS, 'unit': 'TOP/s',
           'frac': None, 'traffic': None, 'kernel_ms': sweep_ms, 'algorithmic_bytes_per_direction': alg_bytes,
           'algorithmic_rate_TBps': alg_bytes / t / 1e12,
           'note': 'planes held in registers and LDS for the whole record (band kernels: a team of workgroups '
                   'per plane trading edge rows every step; column-strip kernels: one workgroup per plane), '
                   'so the 16 B/cell-update algorithmic rate exceeds the HBM peak; VALU lane-ops per step = '
                   'pair-steps x VALU instructions per pair-step of each size\'s adjoint kernel (counter passes), '
                   'traffic = their measured HBM bytes per pair-step x pair-steps'}
    try:
        with open(os.path.join(ROOT, 'profiles', 'sweep_uneven_counters.json')) as f:
            rec = json.load(f)['kernels']
    except (OSError, ValueError, KeyError):
        return out
    names = {(70, 70): 'strip_adjoint_70x70', (100, 140): 'band_adjoint_100x140', (200, 200): 'band_adjoint_200x200'}
    ops = traffic = 0.0
    for a, nb in zip(acqs, rows):
        k = rec.get(names.get((a.nz, a.nx), ''))
        if not k:
            return out
        pairs = nb * a.ns * a.nt * (a.nz + 80) * (a.nx + 80) / 2
        ops += pairs * k['valu_per_pair_step']
        traffic += pairs * (k['hbm_read_bytes'] + k['hbm_write_bytes']) / k['pair_steps']
    achieved = ops / t / 1e12
    out.update(achieved=achieved, frac=achieved / PEAK_VALU_TOPS, traffic=traffic,
               hbm={'achieved': traffic / t / 1e9, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
                    'frac': traffic / t / 1e9 / PEAK_HBM_GBS},
               valu_busy={n: rec[n]['valu_busy'] for n in names.values()},
               source='profiles/sweep_uneven_counters.json')
    return out


def autograd_roofline(kernel, per, alg, pairs):
    """Roofline of the differentiation workload's slower direction: the batched workload's
    column-strip kernels (VALU instructions per pair-step from its counter passes), the
    forward counted twice (the segment recompute)."""
    t = per[kernel] * 1e-3
    out = {'bound': 'valu', 'kernel': kernel, 'achieved': None, 'peak': PEAK_VALU_TOPS, 'unit': 'TOP/s',
           'frac': None, 'traffic': None, 'kernel_ms_per_step': per, 'algorithmic_bytes_per_direction': alg,
           'algorithmic_rate_TBps': alg / t / 1e12,
           'note': 'VALU lane-ops = pair-steps x VALU instructions per pair-step of the batched workload\'s '
                   'kernels (the forward twice: forward and segment recompute); 80 propagations hold at most '
                   '80 of the 256 CUs, which caps the fraction near 0.31.  Kernel times from one extra step with '
                   'events around every launch (recomputes then on the caller stream; the timed steps run them '
                   'on a side stream)'}
    rec = counters()
    if not rec:
        return out
    vp = {'forward': 2 * rec['kernels']['forward']['valu_per_pair_step'],
          'adjoint': rec['kernels']['adjoint']['valu_per_pair_step']}
    achieved = pairs * vp[kernel] / t / 1e12
    out.update(achieved=achieved, frac=achieved / PEAK_VALU_TOPS, source='profiles/rtm_batched_counters.json')
    return out


def median_rate(units, seconds):
    return units / float(np.median(seconds))


class BoardPower:
    """Samples the GPU's board power (hwmon power1_input, microwatts) every 10 ms on a
    host thread while the context is open; `summary()` gives the mean and maximum in
    watts and the board's power cap when the driver exposes them, else None."""

    def __init__(self):
        import glob
        files = sorted(glob.glob('/sys/class/drm/card*/device/hwmon/hwmon*/power1_input'))
        self.path = files[0] if files else None
        self.samples, self._stop, self._thread = [], False, None

    def _read(self, path):
        try:
            with open(path) as f:
                return float(f.read()) / 1e6
        except (OSError, ValueError):
            return None

    def __enter__(self):
        import threading
        if self.path and self._read(self.path) is not None:
            def loop():
                while not self._stop:
                    v = self._read(self.path)
                    if v is not None:
                        self.samples.append(v)
                    time.sleep(0.01)
            self._thread = threading.Thread(target=loop, daemon=True)
            self._thread.start()
        return self

    def __exit__(self, *exc):
        self._stop = True
        if self._thread:
            self._thread.join()

    def summary(self):
        if len(self.samples) < 4:
            return None
        s = np.asarray(self.samples[len(self.samples) // 4:])   # skip the ramp at the start
        cap = self._read(self.path.replace('power1_input', 'power1_cap'))
        return {'mean_W': float(s.mean()), 'max_W': float(s.max()), 'samples': int(len(s)),
                'cap_W': cap, 'source': 'hwmon power1_input sampled every 10 ms during the timed steps'}


# ------------------------------------------------------------------ headline
def workload(n_models, seed, sim):
    """Models, observed gathers (device), backgrounds and mute velocities."""
    from shotsweep_amd import background as BG, synth
    models = synth.mixed(n_models, seed=seed)
    obs = sim.forward(models)[0].contiguous()
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


def protocol_runs(sim, bgs, obs_host, v_mute, runs=3):
    """The images call with host buffers at both ends: observed gathers copied from
    pinned host memory (on a copy stream while the forward stage runs), coefficients
    formed from the host backgrounds, images and illumination copied back to pinned
    host memory, all inside the timed region.
    Returns the wall seconds of each run."""
    g = sim.geom
    obs_pin = torch.empty(obs_host.shape, dtype=torch.float32, pin_memory=True)
    obs_pin.copy_(torch.from_numpy(obs_host))
    nb = len(bgs)
    img_pin = torch.empty(nb, g.ns, g.nz, g.nx, dtype=torch.float32, pin_memory=True)
    ill_pin = torch.empty_like(img_pin)
    vm_pin = torch.tensor(v_mute, dtype=torch.float32).pin_memory()
    vm_dev = torch.empty(nb, dtype=torch.float32, device=sim.kdev)
    out = []
    for _ in range(runs + 1):                    # the first run warms the pinned paths
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vm_dev.copy_(vm_pin, non_blocking=True)
        prep = sim._prepare(bgs)
        img, illum = sim.images_device(prep, obs_pin, 3, vm_dev)  # uploaded beside the forward stage
        img_pin.copy_(img, non_blocking=True)
        ill_pin.copy_(illum, non_blocking=True)
        torch.cuda.synchronize()
        out.append(time.perf_counter() - t0)
    return out[1:], img_pin.numpy(), ill_pin.numpy()


def run_batched(args):
    from shotsweep_amd import gpu
    dev = torch.device('cuda')
    sim = gpu.Sim(device=dev, history=args.history)
    models, obs, obs_host, bgs, v_mute, fallbacks = workload(args.models, args.seed, sim)
    prep = sim._prepare(bgs)
    vm = sim._tensor(v_mute).contiguous()
    nb = prep['nb']
    shot_images = nb * sim.geom.ns
    seg = sim._segment(nb)
    log(f'{args.models} models -> {nb} backgrounds -> {shot_images} shot-images; history segment {seg}; '
        f'surface-velocity fallbacks {fallbacks}')
    for _ in range(args.warmup):
        sim.images_device(prep, obs, 3, vm)
    torch.cuda.synchronize()
    sim.timing, sim.timing_log = True, []
    step_s = []
    t_all = time.perf_counter()
    with BoardPower() as power:
        for _ in range(args.steps):
            t0 = time.perf_counter()
            img, illum = sim.images_device(prep, obs, 3, vm)
            torch.cuda.synchronize()
            step_s.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t_all
    sim.timing = False
    per_kernel = {}
    for rec in sim.timing_log:
        for name, (n, ms) in rec.items():
            if n:
                per_kernel.setdefault(name, []).append(ms / n)
    avg = {k: float(np.mean(v)) for k, v in per_kernel.items()}
    dominant = max(avg, key=avg.get)
    H, W, nt = sim.H, sim.W, sim.geom.nt
    alg_bytes = BYTES_PER_CELL_UPDATE * H * W * nt * shot_images       # one cell-update per cell-step per launch
    ms_per_step = float(np.median(step_s)) * 1e3
    log(f'{ms_per_step:.1f} ms/step (median of {args.steps}), kernels {avg}')
    result = {
        'metric': METRIC, 'value': shot_images / (ms_per_step * 1e-3), 'unit': 'shot-images/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded mixed-structure models)',
        'config': {'workload': 'batched RTM: 64 mixed 70x70 models x 3 backgrounds x 5 shots = 960 shot-images, '
                               'nt=1000, dt=1 ms, dx=10 m, pad 40 (150x150), 15 Hz Ricker',
                   'models': args.models, 'shot_images': shot_images, 'seed': args.seed, 'history_segment': seg},
        'timing': {'value_from': 'median of per-step wall times, inputs resident in HBM',
                   'step_ms': [s * 1e3 for s in step_s], 'mean_rate': shot_images * args.steps / wall},
        'roofline': roofline(dominant, avg[dominant], alg_bytes),
        'kernel_ms': avg,
        'board_power': power.summary(),
    }
    prot_s, pimg, pill = protocol_runs(sim, bgs, obs_host, v_mute)
    result['protocol'] = {
        'value': median_rate(shot_images, prot_s), 'unit': 'shot-images/s', 'runs_ms': [s * 1e3 for s in prot_s],
        'timed': 'observed gathers H2D from pinned host memory (on a copy stream beside the forward stage) + host '
                 'coefficient setup + the native forward and adjoint stages + images and illumination D2H to pinned '
                 'host memory; median of 3 runs after one warm-up',
        'images_equal_device_resident': bool(np.array_equal(pimg, img.cpu().numpy()) and
                                             np.array_equal(pill, illum.cpu().numpy()))}
    log(f'protocol {result["protocol"]["value"]:.0f} shot-images/s')
    if not args.no_cpu:
        rows = 3 * args.cpu_models             # cpu_models models x 3 backgrounds x 5 shots
        base, cimg, cillum, cimg64 = cpu_baseline(bgs, obs_host, v_mute, rows)
        gimg = img[:rows].cpu().numpy()
        err = float(np.linalg.norm(gimg - cimg) / np.linalg.norm(cimg))
        result['cpu_baseline'] = base
        result['speedup_vs_cpu'] = result['value'] / base['value']
        result['image_rel_l2_vs_cpu_oracle'] = err
        result['illum_bitwise_equal_cpu_oracle'] = bool(np.array_equal(illum[:rows].cpu().numpy(), cillum))
        g3 = gimg[:3].astype(np.float64)
        result['image_rel_l2_vs_fp64_oracle'] = float(np.linalg.norm(g3 - cimg64) / np.linalg.norm(cimg64))
        log(f'cpu baseline {base["value"]:.1f} shot-images/s on {base["cores"]} threads')
    return result


# ---------------------------------------------------------------- uneven leg
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


def mixed_kernels(ms, acqs):
    """Kernels each grid size of a mixed sweep runs on: tiled (a tiling was chosen),
    band (a team of workgroups per plane) or column strip (one workgroup per plane)."""
    cus = torch.cuda.get_device_properties(ms.device).multi_processor_count
    out = {}
    for k, a in enumerate(acqs):
        r = ms._round(k)
        if any(t != (0, 0, 0) for t in ms.tiles[k].values()) or r <= 0:
            out[f'{a.nz}x{a.nx}'] = 'tiled'
        elif cus // r > 1:
            out[f'{a.nz}x{a.nx}'] = f'band, {cus // r} workgroups per plane'
        else:
            out[f'{a.nz}x{a.nx}'] = 'column strip, one workgroup per plane'
    return out


def run_mixed(args, steps=None, warmup=None):
    """Uneven batch: 32 models each of 70x70 (nt 1000, 5 shots), 100x140 (nt 1500,
    7 shots) and 200x200 (nt 2000, 10 shots) x 3 backgrounds in one mixed sweep."""
    from shotsweep_amd.mixed import MixedSweep, config_acquisitions
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    acqs = config_acquisitions()
    batches = mixed_workload(args.models_per_size, args.seed, acqs)
    ms = MixedSweep(acqs)
    shot_images = sum(len(b['v_bg']) * a.ns for a, b in zip(acqs, batches))
    rows = [len(b['v_bg']) for b in batches]
    nchunks, _ = ms._chunks(rows, [b['obs_repeat'] for b in batches])
    log(f'mixed: rows {rows}, {shot_images} shot-images, {nchunks} chunks')
    t0 = time.perf_counter()
    for _ in range(max(1, warmup)):
        ms.images_device(batches)
    torch.cuda.synchronize()
    log(f'mixed warmup (with autotuning) {time.perf_counter() - t0:.1f} s, tiles {ms.tiles}')
    step_s = []
    ms.events = []
    for _ in range(steps):
        t0 = time.perf_counter()
        out = ms.images_device(batches)
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - t0)
    events, ms.events = ms.events, None
    per = {}
    for name, a, b in events:
        per[name] = per.get(name, 0.0) + a.elapsed_time(b) / steps
    cell_updates = sum(len(b['v_bg']) * a.ns * a.nt * (a.nz + 80) * (a.nx + 80) for a, b in zip(acqs, batches))
    alg_bytes = BYTES_PER_CELL_UPDATE * cell_updates
    ms_per_step = float(np.median(step_s)) * 1e3
    log(f'mixed {ms_per_step:.1f} ms/step, sweep ms per step {per}')
    result = {
        'metric': 'uneven-batch RTM shot-images/s', 'value': shot_images / (ms_per_step * 1e-3),
        'unit': 'shot-images/s', 'steps': steps, 'warmup': max(1, warmup), 'ms_per_step': ms_per_step,
        'step_ms': [s * 1e3 for s in step_s], 'dtype': 'f32',
        'config': {'workload': 'uneven batched RTM: 32 models each of 70x70 (nt 1000, 5 shots), 100x140 '
                               '(nt 1500, 7 shots), 200x200 (nt 2000, 10 shots) x 3 backgrounds, pad 40, '
                               'dx 10 m, 15 Hz Ricker', 'models_per_size': args.models_per_size,
                   'shot_images': shot_images, 'seed': args.seed, 'chunks': nchunks,
                   'tiles': {f'{a.nz}x{a.nx}': {('forward', 'adjoint')[d]: t for d, t in ms.tiles[k].items()}
                             for k, a in enumerate(acqs)},
                   'kernels': mixed_kernels(ms, acqs)},
        'roofline': dict(sweep_roofline(acqs, rows, per['sweep_adjoint'], alg_bytes), sweep_ms_per_step=per),
    }
    if not args.no_cpu:
        base, imgs = mixed_cpu_baseline(acqs, batches, shot_images)
        result['cpu_baseline'] = base
        result['speedup_vs_cpu'] = result['value'] / base['value']
        errs, ill_eq = [], []
        for (cimg, cill), (gimg, gill) in zip(imgs, out):
            g = gimg[:1].cpu().numpy()
            errs.append(float(np.linalg.norm(g - cimg) / np.linalg.norm(cimg)))
            ill_eq.append(bool(np.array_equal(gill[:1].cpu().numpy(), cill)))
        result['image_rel_l2_vs_cpu_oracle'] = errs
        result['illum_bitwise_equal_cpu_oracle'] = ill_eq
    return result


# -------------------------------------------------------- differentiation leg
def run_autograd(args, steps=None, warmup=None):
    """Differentiation workload: misfit 0.5 * sum((F(v) - obs)^2) and its gradient through
    the autograd Function for 16 models x 5 shots (70 x 70, nt 1000), start = smoothed truth
    (sigma 8), checkpoints every 50 steps with recompute (the reference's scheme)."""
    from shotsweep_amd import autograd, gpu, synth
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    sim = gpu.Sim(device='cuda', checkpoint_interval=50, history='checkpoint')
    true = synth.mixed(16, seed=args.seed)
    obs = sim.forward(true)[0].contiguous()
    start = synth.smooth(true, 8.0)
    v = torch.tensor(start, dtype=torch.float32, device='cuda')
    ns = sim.geom.ns
    for _ in range(max(1, warmup)):
        autograd.misfit_and_gradient(v, obs, sim)
    torch.cuda.synchronize()
    step_s = []
    for _ in range(steps):
        t0 = time.perf_counter()
        loss, grad = autograd.misfit_and_gradient(v, obs, sim)
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - t0)
    # kernel breakdown from one more step with HIP events around every launch (the
    # segment recomputes then run on the caller stream, so each interval is one launch)
    sim.timing, sim.timing_log = True, []
    autograd.misfit_and_gradient(v, obs, sim)
    sim.timing = False
    per = {'forward': sum(t['forward'][1] for t in sim.timing_log),
           'adjoint': sum(t['adjoint'][1] for t in sim.timing_log)}
    dominant = max(per, key=per.get)
    units = len(true) * ns
    alg = BYTES_PER_CELL_UPDATE * sim.H * sim.W * sim.geom.nt * units
    ms_per_step = float(np.median(step_s)) * 1e3
    log(f'autograd {ms_per_step:.2f} ms/step, kernels {per}')
    result = {
        'metric': 'misfit-gradient shot-gradients/s', 'value': units / (ms_per_step * 1e-3),
        'unit': 'shot-gradients/s', 'steps': steps, 'warmup': max(1, warmup), 'ms_per_step': ms_per_step,
        'step_ms': [s * 1e3 for s in step_s], 'dtype': 'f32',
        'config': {'workload': 'differentiation: misfit gradient of 16 mixed 70x70 models x 5 shots, nt 1000, '
                               'autograd.Function, checkpoints every 50 steps with recompute',
                   'models': len(true), 'seed': args.seed, 'misfit': loss},
        'roofline': autograd_roofline(dominant, per, alg, units * sim.geom.nt * sim.H * sim.W / 2),
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
    return result


# ----------------------------------------------------------------- refine leg
def run_refine(args):
    """Learned preconditioner against the plain adjoint-gradient step on a bounded
    sample: args.refine_pairs training pairs, args.refine_models held-out models, 10
    iterations each.  Misfit ratios, iterations for the learned step to reach the
    plain step's final misfit, model RMS errors, and the host-CPU oracle's time for
    one plain iteration of one model (gradient + line-search forwards)."""
    from shotsweep_amd import refine
    t0 = time.perf_counter()
    res, _ = refine.run(pairs=args.refine_pairs, epochs=8, n_eval=args.refine_models, iters=10, seed=args.seed,
                        log=lambda m: None)
    wall = time.perf_counter() - t0
    out = {'metric': 'learned vs plain refine misfit', 'unit': 'mean misfit ratio after k iterations',
           'config': {'pairs': res['pairs'], 'epochs': res['epochs'], 'eval_models': res['eval_models'],
                      'iters': res['iters'], 'seed': args.seed},
           'plain_mean_misfit_ratio': res['plain_mean_misfit_ratio'],
           'learned_mean_misfit_ratio': res['learned_mean_misfit_ratio'],
           'learned_iters_to_reach_plain_final': res['learned_iters_to_reach_plain_final'],
           'learned_fallbacks_per_iter': res['learned_fallbacks_per_iter'],
           'model_rms_error_m_per_s': {k: res[f'model_rms_error_{k}'] for k in ('start', 'plain', 'learned')},
           'openfwi_scores': {k: res[f'openfwi_scores_{k}'] for k in ('start', 'plain', 'learned')},
           'time_s': {'pairs': res['time_pairs_s'], 'train': res['time_train_s'],
                      'refine_plain': res['time_refine_plain_s'], 'refine_learned': res['time_refine_learned_s'],
                      'total': wall},
           'refine_iterations_per_s': res.get('refine_iterations_per_s')}
    if not args.no_cpu:
        from oracle import fd
        from shotsweep_amd import synth
        cores = len(os.sched_getaffinity(0))
        torch.set_num_threads(cores)
        osim = fd.OracleSim()
        true = synth.mixed(1, seed=10_000 + args.seed)
        start = synth.smooth(true, 8.0)
        obs = osim.gathers(true)
        t0 = time.perf_counter()
        g, _ = osim.gradient(start, obs)
        d = -g.sum(1)
        d = d / np.abs(d).max()
        osim.gathers(np.clip(start + 100.0 * d, 1500, 4500))     # one line-search trial
        dt = time.perf_counter() - t0
        from shotsweep_amd import gpu
        gg, _ = gpu.Sim(device='cuda').gradient(start, obs)
        want = g.sum(1).astype(np.float64)
        out['gradient_rel_l2_vs_cpu_oracle'] = float(np.linalg.norm(gg.sum(1) - want) / np.linalg.norm(want))
        out['cpu_baseline'] = {'value': 1.0 / dt, 'unit': 'model-iterations/s', 'cores': cores, 'kind': 'port',
                               'sample': f'one plain iteration of one model (oracle gradient + one trial forward, '
                                         f'FP32 eager PyTorch on {cores} threads), {dt:.1f} s'}
        if out['refine_iterations_per_s']:
            out['speedup_vs_cpu'] = out['refine_iterations_per_s'] / out['cpu_baseline']['value']
    return out


# -------------------------------------------------------------------- probes
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


LEGS = ('uneven', 'differentiation', 'refine')


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
    ap.add_argument('--workload', default='batched', choices=('batched', 'mixed', 'autograd', 'refine'),
                    help='batched: the headline line with its legs; mixed / autograd / refine: one leg alone')
    ap.add_argument('--legs', default='all', help="comma-separated subset of differentiation,uneven,refine, "
                                                  "'all' or 'none' (batched workload only)")
    ap.add_argument('--leg-steps', type=int, default=3, help='timed runs of each leg (median)')
    ap.add_argument('--models-per-size', type=int, default=32)
    ap.add_argument('--refine-pairs', type=int, default=2000)
    ap.add_argument('--refine-models', type=int, default=32)
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
        return print(json.dumps(run_mixed(args)), flush=True)
    if args.workload == 'autograd':
        return print(json.dumps(run_autograd(args)), flush=True)
    if args.workload == 'refine':
        return print(json.dumps(run_refine(args)), flush=True)
    result = run_batched(args)
    legs = LEGS if args.legs == 'all' else () if args.legs == 'none' else tuple(args.legs.split(','))
    result['legs'] = {}
    for leg in legs:
        torch.cuda.empty_cache()     # return the previous workload's buffers to the device
        t0 = time.perf_counter()
        if leg == 'differentiation':
            r = run_autograd(args, steps=args.leg_steps, warmup=1)
        elif leg == 'uneven':
            r = run_mixed(args, steps=args.leg_steps, warmup=1)
        elif leg == 'refine':
            r = run_refine(args)
        else:
            raise SystemExit(f'unknown leg {leg}')
        r['leg_wall_s'] = time.perf_counter() - t0
        result['legs'][leg] = r
        log(f'leg {leg} done in {r["leg_wall_s"]:.1f} s')
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
