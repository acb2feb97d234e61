"""ctypes binding of the C ABI in ``include/pyvibdmc_amd.h``.

The library is loaded from this package directory.  There is no fallback: if the HIP
library is missing every call raises, so nothing silently runs on the CPU.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = 'libpyvibdmc_amd.so'
MAX_COMP = 96

PVD_FP64, PVD_FP32 = 0, 1
PVD_DISCRETE, PVD_CONTINUOUS = 0, 1
PES_EXTERNAL, PES_HARMONIC_X, PES_MORSE_1D, PES_MORSE_BEND, PES_CLUSTER6, PES_PS_WATER = 0, 1, 2, 3, 4, 5
TRIAL_NONE, TRIAL_HARMONIC_1D, TRIAL_GAUSS, TRIAL_HOST = 0, 1, 2, 3

ST_POP, ST_WEIGHT, ST_CAPACITY, ST_BRANCH, ST_SPIN, E_BAD, E_HIP = 1, 3, 4, 5, 6, 7, 8
MASSIVE = "Massive walker birth or death event!!!!!!! Dying..."


class PesDesc(C.Structure):
    _fields_ = [('kind', C.c_int32), ('idx', C.c_int32 * 4), ('p', C.c_double * 12)]


class TrialDesc(C.Structure):
    _fields_ = [('kind', C.c_int32), ('nb', C.c_int32), ('na', C.c_int32),
                ('bi', C.c_int32 * 8), ('bj', C.c_int32 * 8),
                ('ai', C.c_int32 * 4), ('aj', C.c_int32 * 4), ('ak', C.c_int32 * 4),
                ('balpha', C.c_double * 8), ('br0', C.c_double * 8),
                ('aalpha', C.c_double * 4), ('ath0', C.c_double * 4),
                ('c0', C.c_double), ('c1', C.c_double), ('c2', C.c_double)]


class Config(C.Structure):
    _fields_ = [('num_walkers', C.c_int32), ('capacity', C.c_int32), ('n_atoms', C.c_int32),
                ('n_dim', C.c_int32), ('precision', C.c_int32), ('weighting', C.c_int32),
                ('branch_every', C.c_int32), ('num_timesteps', C.c_int32),
                ('delta_t', C.c_double), ('alpha', C.c_double), ('thresh_lower', C.c_double),
                ('sigma', C.c_double * MAX_COMP), ('inv_mass', C.c_double * MAX_COMP),
                ('seed', C.c_uint64), ('pes', PesDesc), ('trial', TrialDesc)]


class DMCError(ValueError):
    """Raised for device-side DMC failures.  Population and weight failures carry the
    reference's message (pyvibdmc.py:397-400, 409-413)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None

_P = C.c_void_p
_SIGS = {
    'pvd_version': (C.c_char_p, []),
    'pvd_last_error': (C.c_char_p, []),
    'pvd_sim_create': (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    'pvd_sim_destroy': (C.c_int, [_P]),
    'pvd_sim_set_walkers': (C.c_int, [_P, _P, C.c_int32, _P, C.c_int64, _P]),
    'pvd_sim_get_walkers': (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_int32), _P]),
    'pvd_sim_start': (C.c_int, [_P, _P]),
    'pvd_sim_start_external': (C.c_int, [_P, _P, _P]),
    'pvd_sim_run': (C.c_int, [_P, C.c_int32, _P]),
    'pvd_sim_run_timed': (C.c_int, [_P, C.c_int32, _P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    'pvd_bench_plane_copy': (C.c_int, [_P, _P, C.c_int32, C.c_int64, C.c_int64, C.c_int32, _P,
                                       C.POINTER(C.c_float)]),
    'pvd_sim_step_injected': (C.c_int, [_P, _P, _P, _P, _P]),
    'pvd_sim_move': (C.c_int, [_P, _P, _P, _P]),
    'pvd_sim_branch_external': (C.c_int, [_P, _P, _P, _P]),
    'pvd_sim_is_cache': (C.c_int, [_P, _P, _P, _P, _P, _P]),
    'pvd_sim_is_propose': (C.c_int, [_P, _P, _P, _P]),
    'pvd_sim_is_accept': (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    'pvd_sim_pes': (C.c_int, [_P, _P, _P]),
    'pvd_sim_planes': (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_int64), C.POINTER(C.c_int32), _P]),
    'pvd_sim_status': (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double), C.POINTER(C.c_int64), _P]),
    'pvd_sim_trajectory': (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, _P, _P, _P]),
    'pvd_sim_desc_begin': (C.c_int, [_P, _P]),
    'pvd_sim_desc_end': (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_int32), _P]),
    'pvd_pes_eval': (C.c_int, [C.POINTER(PesDesc), C.c_int32, C.c_int32, C.c_int32, _P, C.c_int64, _P, _P]),
    'pvd_pes_vjp': (C.c_int, [C.POINTER(PesDesc), C.c_int32, C.c_int32, C.c_int32, _P, C.c_int64, _P, _P, _P]),
    'pvd_trial_eval': (C.c_int, [C.POINTER(TrialDesc), C.c_int32, C.c_int32, C.c_int32, _P, C.c_int64,
                                 _P, _P, _P, _P, _P]),
    'pvd_trial_vjp': (C.c_int, [C.POINTER(TrialDesc), C.c_int32, C.c_int32, C.c_int32, _P, C.c_int64,
                                _P, _P, _P]),
}


def lib_path():
    """The built library: PYVIBDMC_AMD_LIB if set (another build of the same sources),
    else the in-tree build next to this file."""
    return os.environ.get('PYVIBDMC_AMD_LIB') or os.path.join(HERE, LIB_NAME)


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f'{LIB_NAME} is not built ({path}); run `python -m pyvibdmc_amd.build`')
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what=''):
    if rc != 0:
        msg = load().pvd_last_error().decode()
        if rc in (ST_POP, ST_WEIGHT):
            raise DMCError(rc, MASSIVE)
        raise DMCError(rc, f'{what}: {msg}' if what else msg)


def ptr(a):
    """Address of a contiguous NumPy array or torch tensor (or None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags['C_CONTIGUOUS']
        return a.ctypes.data
    return a.data_ptr()


def stream_handle():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ctypes helpers re-exported for the host modules
byref = C.byref
c_int32 = C.c_int32
c_int64 = C.c_int64
c_double = C.c_double


def ctypes_void():
    return C.c_void_p()
