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


C_PROBE = r'''
#include <stddef.h>
#include <stdio.h>
#include "pyvibdmc_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(pvd_config), sizeof(pvd_pes_desc), sizeof(pvd_trial_desc), sizeof(pvd_mlp));
  printf("%zu %zu %zu %zu %zu\n", offsetof(pvd_config, delta_t), offsetof(pvd_config, sigma),
         offsetof(pvd_config, seed), offsetof(pvd_config, pes), offsetof(pvd_config, trial));
  printf("%zu %zu %zu\n", offsetof(pvd_mlp, xmean), offsetof(pvd_mlp, quad), offsetof(pvd_mlp, params));
  printf("%s\n", pvd_version());
  printf("%d\n", pvd_sim_create(NULL, NULL));
  printf("%s\n", pvd_last_error());
  return 0;
}
'''


def test_header_compiles_as_c_and_matches_the_ctypes_layouts(tmp_path):
    """A C11 program built with gcc against include/pyvibdmc_amd.h and linked with the library:
    the header is plain C, the structure sizes and field offsets the C compiler lays out are
    the ones the Python bindings declare, and the entry points that need no GPU answer."""
    import ctypes
    import shutil
    import subprocess
    from pyvibdmc_amd import _native as N
    from pyvibdmc_amd import surrogate as S
    if not os.path.exists(N.lib_path()) or shutil.which('gcc') is None:
        pytest.skip('library not built or no C compiler')
    src, exe = tmp_path / 'probe.c', tmp_path / 'probe'
    src.write_text(C_PROBE)
    libdir = os.path.dirname(N.lib_path())
    subprocess.run(['gcc', '-std=c11', '-Wall', '-Werror', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe),
                    '-L', libdir, '-lpyvibdmc_amd', f'-Wl,-rpath,{libdir}'], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True,
                         env=dict(os.environ, HIP_VISIBLE_DEVICES='')).stdout.splitlines()
    sizes = [int(v) for v in out[0].split()]
    assert sizes == [ctypes.sizeof(N.Config), ctypes.sizeof(N.PesDesc), ctypes.sizeof(N.TrialDesc),
                     ctypes.sizeof(S.MLPDesc)], sizes
    offs = [int(v) for v in out[1].split()]
    assert offs == [N.Config.delta_t.offset, N.Config.sigma.offset, N.Config.seed.offset, N.Config.pes.offset,
                    N.Config.trial.offset], offs
    moffs = [int(v) for v in out[2].split()]
    assert moffs == [S.MLPDesc.xmean.offset, S.MLPDesc.quad.offset, S.MLPDesc.params.offset], moffs
    assert out[3].startswith('pyvibdmc_amd')
    assert int(out[4]) == N.E_BAD and 'null' in out[5]
