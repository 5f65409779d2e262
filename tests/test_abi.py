"""The C-ABI library loads and exports every entry point include/shotsweep.h declares.

No compute calls are made here (no GPU needed): only argument validation,
which fails before any HIP call."""
import ctypes
import re
from pathlib import Path

import pytest

from shotsweep_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / 'include' / 'shotsweep.h').read_text()
    return sorted(set(re.findall(r'\b(ss_\w+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) <= set(names)


def test_version_and_resident_limit():
    lib = _lib.load()
    assert lib.ss_version() == _lib.ABI_VERSION
    assert lib.ss_resident_max_cells() >= 150 * 150


def test_invalid_arguments_are_reported_without_gpu_work():
    lib = _lib.load()
    rc = lib.ss_forward_f32(None, 0, 1, *([None] * 9), None, 1, 0, None)
    assert rc == -1 and b'geometry' in lib.ss_last_error()
    g = _lib.make_geom(1, 5, 70, 70, 40, 1000, 1, (0, 17, 34, 52, 69))
    rc = lib.ss_forward_f32(ctypes.byref(g), 5, 3, *([None] * 9), None, 1, 0, None)
    assert rc == -1 and b'step range' in lib.ss_last_error()
    bad = _lib.make_geom(1, 5, 70, 70, 1, 1000, 1, (0, 17, 34, 52, 69))
    rc = lib.ss_adjoint_f64(ctypes.byref(bad), 0, 1, *([None] * 12), 0, None)
    assert rc == -1 and b'pad >= 2' in lib.ss_last_error()
    g2 = _lib.make_geom(1, 5, 70, 70, 40, 1000, 1, (0, 17, 34, 52, 99))
    rc = lib.ss_image_f32(ctypes.byref(g2), None, None, 1e-4, 1, None, None)
    assert rc == -1 and b'source column' in lib.ss_last_error()


def test_geometry_struct_layout():
    assert ctypes.sizeof(_lib.SSGeom) == 4 * (7 + _lib.MAX_SHOTS)
    with pytest.raises(ValueError):
        _lib.make_geom(1, 2, 70, 70, 40, 10, 1, (0,))
