"""The C-ABI library loads without a GPU and exports every entry point of the header."""
import os
import re

import pytest

from tests.conftest import ROOT


def header_functions():
    src = open(os.path.join(ROOT, 'include', 'pyvibdmc_amd.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|int64_t|const char\*)\s+(pvd_\w+)\s*\(', src, re.M)))


def test_header_lists_entry_points():
    names = header_functions()
    assert 'pvd_sim_run' in names and 'pvd_sim_step_injected' in names and len(names) >= 20


def test_library_exports_every_header_symbol():
    from pyvibdmc_amd import _native as N
    if not os.path.exists(N.lib_path()):
        pytest.skip('library not built')
    lib = N.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    from pyvibdmc_amd import surrogate as S
    bound = set(N.exported_symbols()) | set(S._MSIGS)
    assert set(header_functions()) <= bound, set(header_functions()) - bound
    assert lib.pvd_version().decode().startswith('pyvibdmc_amd')


def test_ctypes_struct_layouts_match_header():
    import ctypes
    from pyvibdmc_amd import _native as N
    # pvd_pes_desc: int32 kind, int32 idx[4], double p[12]
    assert ctypes.sizeof(N.PesDesc) == 4 + 16 + 4 + 96
    assert N.Config.seed.offset % 8 == 0
    assert ctypes.sizeof(N.TrialDesc) % 8 == 0


def test_bad_arguments_are_rejected_without_gpu_work():
    from pyvibdmc_amd import _native as N
    if not os.path.exists(N.lib_path()):
        pytest.skip('library not built')
    lib = N.load()
    assert lib.pvd_sim_create(None, None) == N.E_BAD
    assert b'null' in lib.pvd_last_error()
    cfg = N.Config()
    cfg.num_walkers = 0
    h = N.ctypes_void()
    assert lib.pvd_sim_create(N.byref(cfg), N.byref(h)) == N.E_BAD
