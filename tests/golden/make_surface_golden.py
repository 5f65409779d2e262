"""Write the reference's surface-velocity estimates as golden vectors.

Runs ONLY in the container where the reference is mounted at /root/reference:
it imports the reference's ``openfwi/physics/gpu.py`` and ``background.py`` on
the CPU and records, as data, what ``estimate_surface_velocity`` returns (the
dict, or the ValueError message) for

* the golden flat-layer model's FP32 and FP64 observed gathers (the inputs of
  reference_flat_layers.npz), and
* the five models of the reference's surface-velocity tests: homogeneous
  1500, 2500 and 4500 m/s, and 1500 m/s over 4500 m/s from depth 5 and from
  depth 2 (reference FP32 gathers, ``gathers(models, chunk=3)``),

each on the gathers as they are (FP32, as the reference's tests call it) and
as float64 (as its harness reads gathers, 2_process.py:195), plus the SHA-256
of each gathers array, so a test can check that the
implementation's gathers are the same bytes before it compares estimates.
No closed form pins the estimator's values (its picks, medians, line fits and
quality checks), so these vectors are its parity anchor.

    python tests/golden/make_surface_golden.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
REF = Path('/root/reference/openfwi/physics')


def estimate(bg, obs):
    try:
        e = bg.estimate_surface_velocity(obs)
        return {k: (float(v) if isinstance(v, (float, np.floating)) else bool(v) if isinstance(v, (bool, np.bool_))
                    else v) for k, v in e.items()}
    except ValueError as error:
        return {'error': str(error)}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    sys.path.insert(0, str(REF))
    import gpu            # the reference modules, read-only
    import background as bg
    torch.set_num_threads(8)
    out = {'golden': {}, 'reference_tests': {}}
    with np.load(HERE / 'reference_flat_layers.npz') as g:
        for tag in ('f32', 'f64'):
            obs = g[f'obs_{tag}'][0]
            out['golden'][tag] = {'gathers_sha256': digest(obs), 'estimate': estimate(bg, obs.astype(np.float64))}
    models = [np.full((70, 70), v, dtype=float) for v in (1500, 2500, 4500)]
    for depth in (5, 2):
        m = np.full((70, 70), 1500.0)
        m[depth:] = 4500.0
        models.append(m)
    observed = gpu.Sim(device='cpu').gathers(np.array(models), chunk=3)
    names = ['homogeneous_1500', 'homogeneous_2500', 'homogeneous_4500', 'refractor_depth5', 'refractor_depth2']
    for name, obs in zip(names, observed):
        out['reference_tests'][name] = {'gathers_sha256': digest(obs), 'estimate': estimate(bg, obs),
                                        'estimate_f64': estimate(bg, obs.astype(np.float64))}
    (HERE / 'reference_surface_velocity.json').write_text(json.dumps(out, indent=1, sort_keys=True) + '\n')
    print(json.dumps(out, indent=1))


if __name__ == '__main__':
    main()
