"""ctypes binding of the reference's Partridge-Schwenke water surface, compiled from its
Fortran (sample_potentials/FortPots/Partridge_Schwenke_H2O/calc_h2o_pot.f:1-34, h2opes_v2.f)
by ``oracle/Makefile`` into ``oracle/_ref/libps_ref.so``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  ``calc_hoh_pot(x, nwalk, v)`` takes
x(nwalk, 3, 3) in Fortran order, atoms [H, H, O] in bohr, and returns hartree.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_ref', 'libps_ref.so')
_lib = None


def available():
    return os.path.exists(LIB)


def calc_hoh_pot(cds):
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB)
        _lib.calc_hoh_pot_.restype = None
    x = np.asfortranarray(np.asarray(cds, dtype=np.float64))
    n = ctypes.c_int32(len(x))
    v = np.zeros(len(x))
    _lib.calc_hoh_pot_(x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n), v.ctypes.data_as(ctypes.c_void_p))
    return v
