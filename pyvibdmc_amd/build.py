"""Build recipe for the HIP library (gfx950).

``python -m pyvibdmc_amd.build`` compiles the translation units under ``csrc/`` in
parallel and links them into ``pyvibdmc_amd/libpyvibdmc_amd.so`` in place:
* ``pvd_capi.hip``: the C ABI entry points and auxiliary kernels;
* ``pvd_step_f32.hip`` / ``pvd_step_f64.hip``: the step kernels and their launches for
  float and double walkers;
* ``pvd_mlp.hip``: the surrogate kernels.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
ARCH = os.environ.get('PVD_OFFLOAD_ARCH', 'gfx950')

# -ffp-contract=off: FP64 kernels reproduce NumPy's operation-by-operation rounding.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
FLAGS = ['-O3', f'--offload-arch={ARCH}', '-std=c++17', '-fPIC', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']

TARGETS = [
    ('libpyvibdmc_amd.so', ['pvd_capi.hip', 'pvd_step_f32.hip', 'pvd_step_f64.hip', 'pvd_mlp.hip']),
]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def _filter(text):
    return [l for l in text.splitlines()
            if 'unknown target feature' not in l and 'packed-fp32' not in l and l.strip() != '1 warning generated.']


def _run(cmd, what, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    msgs = _filter(r.stdout)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {what}:\n' + '\n'.join(msgs))
    if verbose and msgs:
        print('\n'.join(msgs), flush=True)


def build(verbose=True, only=None):
    outs = []
    os.makedirs(OBJ, exist_ok=True)
    for lib, srcs in TARGETS:
        if only and lib not in only:
            continue
        src_paths = [os.path.join(CSRC, s) for s in srcs]
        if not all(os.path.exists(p) for p in src_paths):
            continue
        objs = [os.path.join(OBJ, os.path.splitext(s)[0] + '.o') for s in srcs]
        jobs = [([hipcc()] + FLAGS + ['-c', sp, '-o', o], os.path.basename(sp)) for sp, o in zip(src_paths, objs)]
        workers = max(1, min(len(jobs), int(os.environ.get('MAX_JOBS', '0')) or (os.cpu_count() or 1)))
        with ThreadPoolExecutor(max_workers=workers) as ex:
            for f in [ex.submit(_run, c, w, verbose) for c, w in jobs]:
                f.result()
        out = os.path.join(HERE, lib)
        tmp = out + '.tmp'
        _run([hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs, lib, verbose)
        os.replace(tmp, out)
        outs.append(out)
    return outs


# Host AddressSanitizer build of the C ABI layer: pvd_capi.hip's host code instrumented (the
# entry points, argument checks, allocation and clean-up paths), linked with the step and
# surrogate objects of build() and the shared ASan runtime.  Device code is not instrumented.
ASAN_LIB = os.path.join(OBJ, 'libpyvibdmc_amd_asan.so')
ASAN_FLAGS = ['-Xarch_host', '-fsanitize=address', '-Xarch_host', '-fno-omit-frame-pointer']


def asan_runtime():
    """The shared ASan runtime of ROCm's clang (for LD_PRELOAD), or None."""
    import glob
    hits = sorted(glob.glob('/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so'))
    return hits[-1] if hits else None


def build_asan(verbose=True):
    """Build ASAN_LIB (after build(), whose other objects it links); returns its path."""
    others = [os.path.join(OBJ, s) for s in ('pvd_step_f32.o', 'pvd_step_f64.o', 'pvd_mlp.o')]
    if not all(os.path.exists(o) for o in others):
        build(verbose)
    obj = os.path.join(OBJ, 'pvd_capi_asan.o')
    _run([hipcc()] + FLAGS + ASAN_FLAGS + ['-c', os.path.join(CSRC, 'pvd_capi.hip'), '-o', obj], 'pvd_capi.hip (asan)',
         verbose)
    tmp = ASAN_LIB + '.tmp'
    _run([hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC'] + ASAN_FLAGS + ['-shared-libasan', '-o', tmp, obj] +
         others, 'libpyvibdmc_amd_asan.so', verbose)
    os.replace(tmp, ASAN_LIB)
    return ASAN_LIB


if __name__ == '__main__':
    if sys.argv[1:] == ['asan']:
        build_asan()
    else:
        build(only=sys.argv[1:] or None)
