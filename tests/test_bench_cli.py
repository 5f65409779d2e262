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
