"""ss::div12 (the FD kernels' division by 12: reciprocal product plus one FMA
correction, IEEE division when that result is subnormal or not finite) equals
IEEE division bit for bit on all 2^32 float inputs.  The header is compiled for
the host with hipcc, FMA enabled and contraction off."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / 'tests' / 'native' / 'div12_check.cpp'


def test_div12_equals_ieee_division(tmp_path):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not Path(hipcc).exists():
        pytest.skip('hipcc not available')
    exe = tmp_path / 'div12_check'
    subprocess.run([hipcc, '-O2', '-mfma', '-ffp-contract=off', '-pthread',
                    f'-I{ROOT / "shotsweep_amd" / "csrc"}', f'-I{ROOT / "include"}', str(SRC), '-o', str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert 'mismatches=0' in out.stdout
