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


def test_sweep_group_struct_layout():
    # ss_geom (39 int32) + 3 int32 tiling fields, 12 pointers at 8-byte alignment, 4 int32 range fields
    assert ctypes.sizeof(_lib.SSSweepGroup) == 4 * (7 + _lib.MAX_SHOTS) + 12 + 12 * 8 + 16
    assert _lib.SSSweepGroup.c2.offset == 168 and _lib.SSSweepGroup.t0.offset == 264


def test_sweep_validation_without_gpu_work():
    lib = _lib.load()
    arr = (_lib.SSSweepGroup * 5)()
    assert lib.ss_sweep_forward_f32(0, arr, 0, None) == -1
    assert lib.ss_sweep_forward_f32(5, arr, 0, None) == -1
    arr[0].geom = _lib.make_geom(2, 3, 20, 30, 8, 50, 1, (0, 15, 29))
    rc = lib.ss_sweep_forward_f32(1, arr, 0, None)
    assert rc == -1 and b'required' in lib.ss_last_error()
    for f in ('c2', 'A', 'B', 'w', 'amp', 'state', 'hist'):
        setattr(arr[0], f, 64)   # never dereferenced: validation fails first
    rc = lib.ss_sweep_adjoint_f32(1, arr, 0, None)
    assert rc == -1 and b'res, gacc' in lib.ss_last_error()
    arr[0].geom = _lib.make_geom(1, 1, 400, 400, 40, 10, 1, (0,))
    rc = lib.ss_sweep_forward_f32(1, arr, 0, None)
    assert rc == -3 and b'limits' in lib.ss_last_error()
    arr[0].geom = _lib.make_geom(2, 3, 200, 100, 8, 50, 1, (0, 50, 99))
    arr[0].tile_h, arr[0].tile_w, arr[0].depth = 200, 200, 40
    rc = lib.ss_sweep_forward_f32(1, arr, 0, None)
    assert rc == -3 and b'cells per workgroup' in lib.ss_last_error()
    # a forward may keep no Laplacian history (hist NULL); the adjoint needs one
    arr[0].hist = None
    rc = lib.ss_sweep_forward_f32(1, arr, 0, None)
    assert rc == -3 and b'cells per workgroup' in lib.ss_last_error()   # past the pointer checks
    rc = lib.ss_sweep_adjoint_f32(1, arr, 0, None)
    assert rc == -1 and b'hist' in lib.ss_last_error()
    arr[0].hist = 64
    assert lib.ss_sweep_autotune_f32(1, arr, 2, 0, None) == -1
    arr[0].geom = _lib.make_geom(2, 3, 20, 30, 8, 50, 1, (0, 15, 29))
    arr[0].tile_h = arr[0].tile_w = arr[0].depth = 0
    for t0, t1 in ((-1, 10), (10, 10), (20, 51), (50, 0)):
        arr[0].t0, arr[0].t1 = t0, t1
        rc = lib.ss_sweep_forward_f32(1, arr, 0, None)
        assert rc == -1 and b'step range' in lib.ss_last_error(), (t0, t1)


def test_workspace_bytes_closed_form():
    """ss_workspace_bytes (host arithmetic) is the sum of the native call's workspace
    regions, each starting on a 256-byte boundary: per propagation two trace sets
    (gathers, residual), the source term, the imaging sum and 8 state planes; then
    the whole history (nt planes) or 2 ceil(nt/s) checkpoint planes plus two
    segment buffers of s planes; FP32 adds the c2*A plane per row and, for
    segments, the tiled path's coefficient / illumination copies.  The
    reference's own estimate (gpu.py:70-73) is a different scheme."""
    def size(nb, ns, nz, nx, pad, nt, seg, e):
        H, W = nz + 2 * pad, nx + 2 * pad
        nprop, HW = nb * ns, H * W
        regions = ([nb * HW * 4] if e == 4 else []) + [nprop * nt * nx * e] * 2 + [nprop * e, nprop * HW * e,
                                                                                  8 * nprop * HW * e]
        if seg >= nt:
            regions.append(nprop * nt * HW * e)
        else:
            nck = -(-nt // seg)
            regions += [nprop * nck * 2 * HW * e, nprop * seg * HW * e, nprop * seg * HW * e]
            if e == 4:
                regions += [nb * HW * 4, nprop * 4, nprop * 4, nprop * nz * nx * 4, nprop * 4]
        return sum(-(-r // 256) * 256 for r in regions)

    g = _lib.make_geom(192, 5, 70, 70, 40, 1000, 1, (0, 17, 34, 52, 69))
    for seg, e in ((1000, 4), (50, 4), (73, 8), (1000, 8)):
        assert _lib.workspace_bytes(g, seg, e) == size(192, 5, 70, 70, 40, 1000, seg, e), (seg, e)
    assert _lib.workspace_bytes(g, 5000, 4) == _lib.workspace_bytes(g, 1000, 4)
    g2 = _lib.make_geom(3, 7, 11, 13, 5, 37, 1, tuple(range(7)))
    assert _lib.workspace_bytes(g2, 10, 4) == size(3, 7, 11, 13, 5, 37, 10, 4)
    # 86.4 GB of Laplacian history for the 960 shot-images of the batched workload
    assert abs(_lib.workspace_bytes(g, 1000, 4) / 1e9 - 87.7) < 0.2
    with pytest.raises(ValueError, match='segment must be positive'):
        _lib.workspace_bytes(g, 0, 4)
    with pytest.raises(ValueError, match='geometry'):
        _lib.workspace_bytes(_lib.make_geom(0, 5, 70, 70, 40, 1000, 1, (0, 17, 34, 52, 69)), 50, 4)
