"""The C ABI's host code under AddressSanitizer (no GPU needed).

`pyvibdmc_amd.build.build_asan()` compiles pvd_capi.hip with host-side ASan and links it with
the other objects of the library; a child process loads that build with the shared ASan
runtime preloaded and drives the entry points' host paths: the version and error strings,
the size queries, every argument check of pvd_sim_create, and the clean-up after a failed
creation (without a GPU the first device allocation fails, so the partially built
simulation is torn down).  Any heap overflow, use after free or double free aborts the child
with an ASan report.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r'''
import ctypes as C
import sys
sys.path.insert(0, ROOT)
from pyvibdmc_amd import _native as N

lib = N.load()
assert lib.pvd_version()
assert lib.pvd_sim_create(None, None) == N.E_BAD
assert b'null' in lib.pvd_last_error()


def cfg(**kw):
    c = N.Config()
    c.num_walkers, c.capacity, c.n_atoms, c.n_dim = 1000, 0, 3, 3
    c.precision, c.weighting, c.branch_every, c.num_timesteps = N.PVD_FP64, N.PVD_DISCRETE, 1, 10
    c.delta_t, c.alpha, c.thresh_lower, c.seed = 5.0, 0.1, 0.01, 7
    for p in range(9):
        c.sigma[p], c.inv_mass[p] = 0.05, 1.0 / 1837.15
    for k, v in kw.items():
        setattr(c, k, v)
    return c


sim = C.c_void_p()
for bad in (dict(num_walkers=0), dict(n_atoms=0), dict(n_dim=0), dict(n_atoms=N.MAX_COMP),
            dict(precision=5), dict(weighting=9), dict(capacity=1 << 30),
            dict(precision=N.PVD_FP32, capacity=(1 << 30) - 1)):
    rc = lib.pvd_sim_create(C.byref(cfg(**bad)), C.byref(sim))
    assert rc == N.E_BAD, (bad, rc, lib.pvd_last_error())
    assert lib.pvd_last_error()
# well-formed: fails at the first device allocation here, and frees what it had
for w in (N.PVD_DISCRETE, N.PVD_CONTINUOUS):
    for prec in (N.PVD_FP64, N.PVD_FP32):
        rc = lib.pvd_sim_create(C.byref(cfg(weighting=w, precision=prec)), C.byref(sim))
        assert rc != 0 and lib.pvd_last_error(), rc
assert lib.pvd_sim_destroy(None) in (0, N.E_BAD)
for name, args in (('pvd_mlp_param_count', [C.c_int32, C.c_int32]), ('pvd_mlp_pack_bytes', [C.c_int32]),
                   ('pvd_mlp_workspace_bytes', [C.c_int32, C.c_int32, C.c_int64]),
                   ('pvd_mlp_noalloc_scratch_bytes', [])):
    getattr(lib, name).restype = C.c_int64
    getattr(lib, name).argtypes = args
for D, H in ((3, 128), (15, 256)):
    assert lib.pvd_mlp_param_count(D, H) > 0
    assert lib.pvd_mlp_pack_bytes(H) > 0
    assert lib.pvd_mlp_workspace_bytes(D, H, 8192) > 0
assert lib.pvd_mlp_noalloc_scratch_bytes() > 0
maps = open('/proc/self/maps').read()
assert 'libclang_rt.asan' in maps and 'libpyvibdmc_amd_asan' in maps
print('asan driver ok', flush=True)
'''

# a planted overflow in a host-only library built the same way: the harness must report it
PLANT_C = r'''
#include <stdlib.h>
int overflow(void) {
  volatile int* p = (int*)malloc(4 * sizeof(int));
  p[4] = 1;
  int v = p[4];
  free((void*)p);
  return v;
}
'''


def _env():
    from pyvibdmc_amd import build as B
    rt = B.asan_runtime()
    if rt is None:
        pytest.skip("ROCm's clang ASan runtime is not installed")
    if not os.path.exists(B.ASAN_LIB):
        B.build_asan(verbose=False)
    return dict(os.environ, LD_PRELOAD=rt, PYVIBDMC_AMD_LIB=B.ASAN_LIB, ASAN_OPTIONS='detect_leaks=0',
                HIP_VISIBLE_DEVICES='')


def test_the_sanitizer_harness_reports_a_planted_overflow(tmp_path):
    clang = '/opt/rocm/llvm/bin/clang'
    if not os.path.exists(clang):
        pytest.skip("ROCm's clang is not installed")
    env = _env()
    src, so = tmp_path / 'plant.c', tmp_path / 'plant.so'
    src.write_text(PLANT_C)
    subprocess.run([clang, '-shared', '-fPIC', '-g', '-fsanitize=address', '-shared-libasan', '-o', str(so), str(src)],
                   check=True)
    r = subprocess.run([sys.executable, '-c', f'import ctypes; ctypes.CDLL({str(so)!r}).overflow()'], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and 'heap-buffer-overflow' in r.stderr, r.stderr[-2000:]


def test_c_abi_host_paths_under_address_sanitizer():
    env = _env()
    r = subprocess.run([sys.executable, '-c', f'ROOT = {ROOT!r}\n' + DRIVER], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and 'AddressSanitizer' not in r.stderr, r.stdout[-2000:] + r.stderr[-4000:]
    assert 'asan driver ok' in r.stdout
