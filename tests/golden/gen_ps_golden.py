"""Writes tests/golden/ps_water.npz: energies of the reference's Partridge-Schwenke surface
(its own Fortran, compiled by ``make -C oracle``) at 1000 [H, H, O] geometries in bohr.

Run in the development container: python tests/golden/gen_ps_golden.py
Geometries: the surface's equilibrium plus Gaussian noise of 0.05, 0.15 and 0.3 bohr
(600 / 300 / 100 walkers, seed 0), so the fixture covers the region a DMC ensemble visits
and its tails.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pes as P  # noqa: E402
from oracle import ps_ref  # noqa: E402


def main():
    rs = np.random.RandomState(0)
    eq = P.ps_eq()
    cds = np.concatenate([eq[None] + rs.normal(0, s, (n, 3, 3)) for s, n in ((0.05, 600), (0.15, 300), (0.3, 100))])
    v = ps_ref.calc_hoh_pot(cds)
    np.savez(os.path.join(ROOT, 'tests', 'golden', 'ps_water.npz'), cds=cds, v=v)
    print('wrote', len(v), 'energies; range', v.min(), v.max())


if __name__ == '__main__':
    main()
