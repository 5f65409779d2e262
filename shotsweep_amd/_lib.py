"""ctypes binding of libshotsweep.so (declarations in include/shotsweep.h).

The library is loaded from this package directory only; there is no fallback.
A missing or stale library raises RuntimeError telling how to build it.
"""
import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / 'libshotsweep.so'
MAX_SHOTS = 32
IMPL_AUTO, IMPL_STEPWISE, IMPL_RESIDENT, IMPL_TILED, IMPL_UNTILED = 0, 1, 2, 3, 4
RTM_FORWARD, RTM_ADJOINT = 1, 2
ABI_VERSION = 8

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


class SSGeom(ctypes.Structure):
    """Mirror of ss_geom."""
    _fields_ = [('nb', ctypes.c_int32), ('ns', ctypes.c_int32), ('nz', ctypes.c_int32), ('nx', ctypes.c_int32),
                ('pad', ctypes.c_int32), ('nt', ctypes.c_int32), ('row', ctypes.c_int32),
                ('src', ctypes.c_int32 * MAX_SHOTS)]


def make_geom(nb, ns, nz, nx, pad, nt, row, src):
    if len(src) != ns or ns > MAX_SHOTS:
        raise ValueError(f'need one source column per shot (ns <= {MAX_SHOTS})')
    g = SSGeom(nb, ns, nz, nx, pad, nt, row)
    for k, c in enumerate(src):
        g.src[k] = int(c)
    return g


# name -> argtypes (all functions return int unless listed in _RESTYPES)
def _sig(real):
    R = real  # scalar type of the working precision for by-value reals
    geom = ctypes.POINTER(SSGeom)
    return {
        'forward': [geom, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
        'adjoint': [geom, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
        'residual': [geom, _vp, _vp, _i, _vp, _d, _d, _d, _d, _vp, _vp],
        'mute_weights': [geom, _vp, _d, _d, _d, _d, _vp, _vp],
        'gradient': [geom, _vp, _vp, _vp, R, _vp, _vp],
        'image': [geom, _vp, _vp, R, _i, _vp, _vp],
        'born': [geom, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        'lap': [_i64, _i, _i, _vp, _vp, _vp],
        'lap_adjoint': [_i64, _i, _i, _vp, _vp, _vp],
        'fold_edge_padding': [_i64, _i, _i, _i, _vp, _vp, _vp],
    }


class SSSweepGroup(ctypes.Structure):
    """Mirror of ss_sweep_group (one grid-size group of a mixed-size sweep)."""
    _fields_ = [('geom', SSGeom), ('tile_h', ctypes.c_int32), ('tile_w', ctypes.c_int32), ('depth', ctypes.c_int32),
                ('c2', _vp), ('A', _vp), ('B', _vp), ('w', _vp), ('amp', _vp), ('res', _vp), ('state', _vp),
                ('gathers', _vp), ('hist', _vp), ('gacc', _vp), ('illum', _vp), ('src_grad', _vp),
                ('t0', ctypes.c_int32), ('t1', ctypes.c_int32), ('resume', ctypes.c_int32), ('phase', ctypes.c_int32)]


class SSRtmArgs(ctypes.Structure):
    """Mirror of ss_rtm_args (one gradient / images call of the native driver)."""
    _fields_ = [('stages', ctypes.c_int32), ('segment', ctypes.c_int32), ('impl', ctypes.c_int32),
                ('obs_repeat', ctypes.c_int32), ('laplacian', ctypes.c_int32), ('timing', ctypes.c_int32),
                ('dt', _d), ('dx', _d), ('t_mute', _d), ('ramp', _d), ('eps', _d), ('k', _d),
                ('c2', _vp), ('A', _vp), ('B', _vp), ('w', _vp), ('amp', _vp), ('sd', _vp), ('v', _vp),
                ('obs', _vp), ('v_mute', _vp), ('seed', _vp), ('gathers', _vp), ('grad', _vp), ('illum', _vp),
                ('img', _vp)]


SWEEP_MAX_GROUPS = 4
_F = ctypes.c_float
# FP32-only entry points of the learned preconditioner
_CNN = {
    'ss_conv3x3_f32': [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp],
    'ss_conv3x3_dgrad_f32': [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp],
    'ss_conv3x3_wgrad_f32': [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp],
    'ss_relu_grad_f32': [_i64, _vp, _vp, _vp, _vp],
    'ss_mse_f32': [_i64, _vp, _vp, _vp, _vp, _i, _vp],
    'ss_adam_f32': [_i64, _vp, _vp, _vp, _vp, _F, _F, _F, _F, _i, _vp],
    'ss_precond_features_f32': [_i, _i, _i, _vp, _vp, _vp, _vp, _vp],
    'ss_sumsq_diff_f32': [_i, _i64, _vp, _vp, _vp, _vp],
    'ss_velocity_update_f32': [_i, _i64, _vp, _vp, _vp, _F, _F, _vp, _vp],
}

_SWEEP = {
    'ss_sweep_forward_f32': [_i, ctypes.POINTER(SSSweepGroup), _i, _vp],
    'ss_sweep_adjoint_f32': [_i, ctypes.POINTER(SSSweepGroup), _i, _vp],
    'ss_sweep_autotune_f32': [_i, ctypes.POINTER(SSSweepGroup), _i, _i, _vp],
    'ss_sweep_round_f32': [ctypes.POINTER(SSGeom)],
    'ss_sweep_band_errors': [ctypes.POINTER(_i)],
}

_RTM = {
    'ss_rtm_f32': [ctypes.POINTER(SSGeom), ctypes.POINTER(SSRtmArgs), _vp, _i64, _vp],
    'ss_rtm_f64': [ctypes.POINTER(SSGeom), ctypes.POINTER(SSRtmArgs), _vp, _i64, _vp],
    'ss_rtm_noalloc_f32': [ctypes.POINTER(SSGeom), ctypes.POINTER(SSRtmArgs), _vp],
    'ss_rtm_timing': [ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_i), ctypes.POINTER(_d)],
}

_QC = {
    'ss_first_breaks_f32': [_i64, _i, _i, _vp, _d, _vp, _vp, _vp],
    'ss_first_breaks_f64': [_i64, _i, _i, _vp, _d, _vp, _vp, _vp],
    'ss_velocity_scores_f32': [_i, _i, _i, _vp, _vp, _d, _d, _vp, _vp, _vp],
    'ss_velocity_scores_f64': [_i, _i, _i, _vp, _vp, _d, _d, _vp, _vp, _vp],
}

EXPORTS = ['ss_last_error', 'ss_version', 'ss_resident_max_cells', 'ss_workspace_bytes'] + [
    f'ss_{name}_{sfx}' for sfx in ('f32', 'f64') for name in _sig(ctypes.c_float)] + list(_CNN) + list(_SWEEP) + list(_QC) + list(_RTM)

_LIB = None


def workspace_bytes(geom, segment, elem_size):
    """ss_workspace_bytes: device bytes of one gradient/images call (host arithmetic only)."""
    n = load().ss_workspace_bytes(ctypes.byref(geom), int(segment), int(elem_size))
    if n < 0:
        raise ValueError(load().ss_last_error().decode())
    return int(n)


def load():
    """Load the library once; raise RuntimeError when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f'{LIB_PATH} is not built; run `python -m shotsweep_amd.build` '
                           '(shotsweep_amd has no CPU fallback)')
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.ss_last_error.restype = ctypes.c_char_p
    lib.ss_last_error.argtypes = []
    lib.ss_version.restype = ctypes.c_int
    lib.ss_resident_max_cells.restype = ctypes.c_int64
    lib.ss_workspace_bytes.restype = ctypes.c_int64
    lib.ss_workspace_bytes.argtypes = [ctypes.POINTER(SSGeom), ctypes.c_int, ctypes.c_int]
    for sfx, real in (('f32', ctypes.c_float), ('f64', ctypes.c_double)):
        for name, args in _sig(real).items():
            fn = getattr(lib, f'ss_{name}_{sfx}')
            fn.argtypes = args
            fn.restype = ctypes.c_int
    for name, args in list(_CNN.items()) + list(_SWEEP.items()) + list(_QC.items()) + list(_RTM.items()):
        fn_ = getattr(lib, name)
        fn_.argtypes = args
        fn_.restype = ctypes.c_int
    if lib.ss_version() != ABI_VERSION:
        raise RuntimeError(f'libshotsweep ABI {lib.ss_version()} != expected {ABI_VERSION}; rebuild it')
    _LIB = lib
    return lib


def fn(name, dtype):
    """Entry point `ss_<name>_<f32|f64>` for a torch dtype."""
    sfx = {torch.float32: 'f32', torch.float64: 'f64'}.get(dtype)
    if sfx is None:
        raise TypeError(f'shotsweep kernels support float32 and float64, not {dtype}')
    return getattr(load(), f'ss_{name}_{sfx}')


def rtm_timing():
    """{'forward': (launches, ms), 'adjoint': (launches, ms)} of this thread's last
    ss_rtm_* call made with timing = 1 (waits for those launches)."""
    nf, na = ctypes.c_int(), ctypes.c_int()
    mf, ma = ctypes.c_double(), ctypes.c_double()
    check(load().ss_rtm_timing(ctypes.byref(nf), ctypes.byref(mf), ctypes.byref(na), ctypes.byref(ma)), 'rtm timing')
    return {'forward': (nf.value, mf.value), 'adjoint': (na.value, ma.value)}


def check(rc, what=''):
    if rc != 0:
        msg = load().ss_last_error().decode(errors='replace')
        raise RuntimeError(f'shotsweep {what} failed ({rc}): {msg}')


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError('shotsweep kernels take device tensors')
    if not t.is_contiguous():
        raise ValueError('shotsweep kernels take contiguous tensors')
    return t.data_ptr()


def stream_handle(device):
    return torch.cuda.current_stream(device).cuda_stream
