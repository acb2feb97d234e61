"""Write golden descriptor vectors by running the reference's DistIt in this container.

Usage (development container only; the reference does not exist on GPU hosts):

    python tests/golden/gen_distit_golden.py [--ref /root/reference]

The descriptor's permutation sorting (which atom lands at which position, in which order
ties and groups resolve) is not fixed by the invariances the reference's own tests assert,
so its conventions are pinned by the reference's outputs.  The script loads the single
module ``pyvibdmc/simulation_utilities/tensorflow_descriptors/distance_descriptors.py``
from its file (it imports only itertools and NumPy; bytecode writing is off), runs
``DistIt(zs, 'coulomb', sorted_atoms=..., sorted_groups=..., force_numpy=True).run(cds)``
on seeded geometries, and saves inputs and outputs to ``tests/golden/distit.npz``.  The
trimer case uses the coordinate fixtures of the reference's test
(pyvibdmc/tests/test_cupy_dist.py:5-23), read from that file's text as literals, and the
permuted copies its ``test_sorting`` builds (:80-107).
"""
import argparse
import ast
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def load_distit(ref):
    path = os.path.join(ref, 'pyvibdmc', 'simulation_utilities', 'tensorflow_descriptors', 'distance_descriptors.py')
    sys.dont_write_bytecode = True
    spec = importlib.util.spec_from_file_location('_ref_distance_descriptors', path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.DistIt


def test_fixtures(ref):
    """The udu / uuu coordinate literals of the reference's DistIt test."""
    src = open(os.path.join(ref, 'pyvibdmc', 'tests', 'test_cupy_dist.py')).read()
    out = {}
    for node in ast.parse(src).body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name) and \
                node.targets[0].id in ('udu', 'uuu'):
            out[node.targets[0].id] = np.array(ast.literal_eval(node.value.args[0]), np.float64)
    return out['udu'], out['uuu']


def swapped_copies(x):
    """test_sorting's geometries: x, monomers 1 and 2 swapped, H1/H2 of water 1 swapped, both."""
    idx = np.arange(len(x))
    idx[[0, 1, 2]], idx[[3, 4, 5]] = [3, 4, 5], [0, 1, 2]
    w2w1 = x[idx].copy()
    a = x.copy()
    a[[1, 2]] = a[[2, 1]]
    b = w2w1.copy()
    b[[1, 2]] = b[[2, 1]]
    return np.array([x, w2w1, a, b])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--ref', default='/root/reference')
    args = ap.parse_args()
    from oracle import pes as P
    DistIt = load_distit(args.ref)
    rs = np.random.RandomState(2024)
    water = P.water_eq()[None] + rs.normal(0, 0.1, (64, 3, 3))
    cluster = P.cluster6_eq()[None] + rs.normal(0, 0.1, (64, 6, 3))
    udu, uuu = test_fixtures(args.ref)
    trimer = np.concatenate([swapped_copies(udu), swapped_copies(uuu)])
    mono = [[0], [1, 2], [3], [4, 5]]
    cases = {
        'water_like': (water, [1, 1, 8], [[0, 1], [2]], None),
        'cluster_like': (cluster, [8, 1, 1, 8, 1, 1], [[0, 3], [1, 2, 4, 5]], None),
        'cluster_monomers': (cluster, [8, 1, 1, 8, 1, 1], mono, [[0, 1, 2], [3, 4, 5]]),
        'cluster_groups_only': (cluster, [8, 1, 1, 8, 1, 1], None, [[0, 1, 2], [3, 4, 5]]),
        'cluster_reversed_sublist': (cluster, [8, 1, 1, 8, 1, 1], [[3, 0], [5, 4, 2, 1]], None),
        'trimer_reference_fixture': (trimer, [8, 1, 1] * 3, [[0], [1, 2], [3], [4, 5], [6], [7, 8]],
                                     [[0, 1, 2], [3, 4, 5], [6, 7, 8]]),
    }
    out = {}
    for name, (cds, zs, sa, sg) in cases.items():
        t = DistIt(zs, 'coulomb', sorted_atoms=sa, sorted_groups=sg, force_numpy=True)
        out[name + '/cds'] = cds
        out[name + '/zs'] = np.asarray(zs, np.int64)
        out[name + '/spec'] = np.array(json.dumps({'sorted_atoms': sa, 'sorted_groups': sg}))
        out[name + '/coulomb'] = np.asarray(t.run(cds), np.float64)
    np.savez_compressed(os.path.join(HERE, 'distit.npz'), **out)
    print('wrote', os.path.join(HERE, 'distit.npz'), sorted(cases))


if __name__ == '__main__':
    main()
