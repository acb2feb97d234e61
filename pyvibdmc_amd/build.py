"""Build recipe for the HIP library (gfx950).

``python -m pyvibdmc_amd.build`` compiles ``csrc/pvd_capi.hip`` (step kernels and the
simulation C ABI) and ``csrc/pvd_mlp.hip`` (surrogate kernels) into
``pyvibdmc_amd/libpyvibdmc_amd.so`` in place.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
ARCH = os.environ.get('PVD_OFFLOAD_ARCH', 'gfx950')

# -ffp-contract=off: FP64 kernels reproduce NumPy's operation-by-operation rounding.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
FLAGS = ['-O3', f'--offload-arch={ARCH}', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']

TARGETS = [
    ('libpyvibdmc_amd.so', ['pvd_capi.hip', 'pvd_mlp.hip']),
]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose=True, only=None):
    outs = []
    for lib, srcs in TARGETS:
        if only and lib not in only:
            continue
        src_paths = [os.path.join(CSRC, s) for s in srcs]
        if not all(os.path.exists(p) for p in src_paths):
            continue
        out = os.path.join(HERE, lib)
        tmp = out + '.tmp'
        cmd = [hipcc()] + FLAGS + ['-o', tmp] + src_paths
        if verbose:
            print(' '.join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        msgs = [l for l in r.stdout.splitlines()
                if 'unknown target feature' not in l and 'packed-fp32' not in l and l.strip() != '1 warning generated.']
        if r.returncode != 0:
            raise RuntimeError(f'hipcc failed for {lib}:\n' + '\n'.join(msgs))
        if verbose and msgs:
            print('\n'.join(msgs))
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == '__main__':
    build(only=sys.argv[1:] or None)
