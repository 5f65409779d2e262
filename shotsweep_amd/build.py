"""Compile libshotsweep.so (gfx950) in-tree with hipcc.

    python -m shotsweep_amd.build [--verbose]

Flags that are part of the numerical contract:
  -ffp-contract=off    keep every product/sum that the reference rounds
                       separately as a separate rounding (no implicit FMA)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                       no packed FP32 instructions
"""
import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / 'csrc'
LIB = HERE / 'libshotsweep.so'
ARCH = os.environ.get('SHOTSWEEP_ARCH', 'gfx950')

FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', f'--offload-arch={ARCH}', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument', f'-I{ROOT / "include"}']


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not Path(path).exists():
        raise RuntimeError('hipcc not found: ROCm is required to build libshotsweep.so')
    return path


def sources():
    return sorted(CSRC.glob('*.hip'))


def _compile(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False, extra=(), jobs=None):
    """Build the shared library; returns its path.  Always recompiles: every
    source is compiled to an object in parallel (each keeps its own device code;
    no kernel is launched across translation units), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    import tempfile
    tmp = LIB.with_suffix('.so.tmp')
    cflags = [f for f in FLAGS if f != '-shared']
    with tempfile.TemporaryDirectory(prefix='shotsweep_build_') as d:
        objs = [Path(d) / (src.stem + '.o') for src in sources()]
        cmds = [[hipcc(), *cflags, *extra, '-c', str(src), '-o', str(obj)] for src, obj in zip(sources(), objs)]
        workers = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
        with ThreadPoolExecutor(workers) as pool:
            procs = list(pool.map(lambda c: _compile(c, verbose), cmds))
        for proc in procs:
            if proc.returncode != 0:
                raise RuntimeError(f'hipcc failed ({proc.returncode}):\n{proc.stderr[-8000:]}')
            if verbose and proc.stderr.strip():
                print(proc.stderr[-4000:], file=sys.stderr)
        proc = _compile([hipcc(), *FLAGS, *extra, *map(str, objs), '-o', str(tmp)], verbose)
        if proc.returncode != 0:
            raise RuntimeError(f'hipcc link failed ({proc.returncode}):\n{proc.stderr[-8000:]}')
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--verbose', action='store_true')
    args = ap.parse_args()
    print(build(args.verbose))


if __name__ == '__main__':
    main()
