"""Pins of the oracle's Partridge-Schwenke surface (oracle.pes.ps_water) against the
reference's own Fortran: the committed energies it wrote (tests/golden/ps_water.npz, by
tests/golden/gen_ps_golden.py) and, where the compiled library is present (oracle/_ref), fresh
random geometries; plus the surface's symmetry and minimum."""
import numpy as np
import pytest

from oracle import pes as P
from oracle import ps_ref
from tests.conftest import GOLDEN


def test_oracle_matches_reference_energies():
    g = np.load(GOLDEN + '/ps_water.npz')
    np.testing.assert_allclose(P.ps_water(g['cds']), g['v'], rtol=1e-12, atol=1e-15)


@pytest.mark.skipif(not ps_ref.available(), reason='oracle/_ref not built (make -C oracle)')
def test_oracle_matches_the_compiled_reference():
    x = P.ps_eq()[None] + np.random.RandomState(11).normal(0, 0.2, (5000, 3, 3))
    np.testing.assert_allclose(P.ps_water(x), ps_ref.calc_hoh_pot(x), rtol=1e-12, atol=1e-15)


def test_symmetry_and_minimum():
    x = P.ps_eq()[None] + np.random.RandomState(12).normal(0, 0.1, (500, 3, 3))
    np.testing.assert_allclose(P.ps_water(x[:, [1, 0, 2]]), P.ps_water(x), rtol=1e-12, atol=1e-15)
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    np.testing.assert_allclose(P.ps_water(x @ R.T + 2.0), P.ps_water(x), rtol=1e-9, atol=1e-14)
    eq = P.ps_eq()[None]
    assert abs(P.ps_water(eq)[0]) < 1e-6  # the published minimum, 0 to the fit's accuracy
    assert np.all(P.ps_water(x) > P.ps_water(eq)[0])
