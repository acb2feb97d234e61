"""Harmonic analysis and initial ensembles (simulation_utilities/initial_conditioner/).

* The oracle's element-by-element Hessian (harmonic_analysis.py:54-91) gives the Wilson
  GF-matrix harmonic frequencies of the triatomic Morse+bend surface (force constants
  2 De a^2 for each stretch, kb for the bend) to 0.02 cm^-1, and six zero modes.
* The package's batched Hessian (all stencils in one getpot call) equals the oracle's bit
  for bit.
* Normal-mode sampling: the mass-weighted normal coordinates of the ensemble have variance
  1 / omega (the harmonic ground state psi_0 as a walker density); like-atom permutation
  keeps each walker's atoms.
"""
import numpy as np
import pytest

import pyvibdmc_amd as pv
from oracle import initial as OI
from oracle import pes as P

WN = 4.556335281212229e-6


def eq_water():
    return P.water_eq()


def test_oracle_hessian_gives_gf_frequencies():
    h = OI.hessian(eq_water(), P.morse_bend)
    md = np.repeat([P.M_H, P.M_H, P.M_O], 3)
    lam = np.linalg.eigvalsh(h / np.sqrt(md) / np.sqrt(md[:, None]))
    f = np.sqrt(np.abs(lam)) * np.sign(lam) / WN
    gf = OI.gf_frequencies(P.M_H, P.M_O, P.MB_RE, P.MB_THETA0, 2 * P.MB_DE * P.MB_A ** 2, P.MB_KB) / WN
    np.testing.assert_allclose(np.sort(f)[6:], gf, atol=0.02)
    assert np.all(np.sort(np.abs(f))[:6] < 5.0)  # translations and rotations, up to FD noise
    # the coupled stretches split around the local-mode frequency a sqrt(2 De / mu) = 3921.7
    assert abs(0.5 * (gf[1] + gf[2]) - 3921.7) < 2.0 and abs(gf[0] - 1662.2) < 10


def test_batched_hessian_equals_elementwise():
    ha = pv.HarmonicAnalysis(eq_water(), ['H', 'H', 'O'], pv.Potential_Direct(P.morse_bend),
                             masses=[P.M_H, P.M_H, P.M_O])
    np.testing.assert_array_equal(ha.generate_hessian(), OI.hessian(eq_water(), P.morse_bend))
    freqs, modes = ha.run()
    assert freqs.shape == (9,) and modes.shape == (9, 9)
    np.testing.assert_allclose(modes.T @ modes, np.eye(9), atol=1e-12)


def test_harmonic_sampling_has_ground_state_variance():
    ha = pv.HarmonicAnalysis(eq_water(), ['H', 'H', 'O'], pv.Potential_Direct(P.morse_bend),
                             masses=[P.M_H, P.M_H, P.M_O])
    freqs, modes = ha.run()
    np.random.seed(3)
    n = 200_000
    ic = pv.InitialConditioner(coord=eq_water(), atoms=['H', 'H', 'O'], num_walkers=n,
                               technique='harmonic_sampling', masses=np.array([P.M_H, P.M_H, P.M_O]),
                               technique_kwargs=dict(freqs=freqs, normal_modes=modes, scaling_factor=1.0,
                                                     ensemble=None))
    x = ic.run()
    assert x.shape == (n, 3, 3)
    md = np.repeat([P.M_H, P.M_H, P.M_O], 3)
    q = ((x - eq_water()).reshape(n, 9) * np.sqrt(md)) @ modes[:, 6:]
    w = freqs[6:] * WN
    np.testing.assert_allclose(q.var(axis=0) * w, 1.0, rtol=0.01)
    assert np.all(np.abs(q.mean(axis=0)) < 5 * np.sqrt(1 / w / n))
    # walkers sample psi_0 itself (variance 1/omega per mode, the DMC walker density), so their
    # mean harmonic potential is sum omega / 2, the harmonic ZPE, up to the Morse anharmonicity
    v = P.morse_bend(x[:20000])
    assert abs(v.mean() / WN - 0.5 * freqs[6:].sum()) < 0.05 * 0.5 * freqs[6:].sum()


def test_permute_like_atoms():
    ch5 = np.array([[1.0, 2.0, 3.0], [0.13, 2.09, 0.0], [1.79, -1.39, 0.0], [2.23, 0.36, 0.0],
                    [-0.82, -0.63, -1.78], [-0.82, -0.63, 1.78]])[None]
    np.random.seed(0)
    ic = pv.InitialConditioner(coord=ch5, atoms=['C', 'H', 'H', 'H', 'H', 'H'], num_walkers=500,
                               technique='permute_atoms',
                               technique_kwargs={'like_atoms': [[0], [1, 2, 3, 4, 5]], 'ensemble': None})
    w = ic.run()
    assert w.shape == (500, 6, 3)
    np.testing.assert_array_equal(w[:, 0], np.repeat(ch5[:, 0], 500, axis=0))
    ref = sorted(map(tuple, ch5[0, 1:]))
    for k in range(500):
        assert sorted(map(tuple, w[k, 1:])) == ref
    assert len({tuple(w[k, 1]) for k in range(500)}) == 5  # the first H slot takes every H


def test_dipole_derivatives_of_a_linear_dipole():
    q = np.array([0.4, 0.4, -0.8])

    class Dip:
        @staticmethod
        def getpot(c):
            return (q[None, :, None] * c).sum(axis=1)
    ha = pv.HarmonicAnalysis(eq_water(), ['H', 'H', 'O'], None, masses=[P.M_H, P.M_H, P.M_O], dipole=Dip())
    d = ha.dipole_derivs()
    np.testing.assert_allclose(d, np.kron(q[:, None], np.eye(3)), atol=1e-9)
