"""CPU restatement of the reference's harmonic analysis, element by element.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Follows
``simulation_utilities/initial_conditioner/harmonic_analysis.py:54-91`` (one ``getpot``
call per Hessian element, in the reference's loop order) and ``finite_difference.py``.
Pinned by the Wilson GF-matrix harmonic frequencies of the triatomic Morse+bend surface
(``gf_frequencies``, a textbook closed form) in ``tests/test_initial_conditions.py``.
"""
import itertools as itt

import numpy as np

W1 = {3: np.array([-1 / 2, 0, 1 / 2]), 5: np.array([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12])}
W2 = {3: np.array([1, -2, 1]), 5: np.array([-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])}


def stencil(eq, atm_cd, dx, npts):
    steps = np.arange(-(npts // 2), npts // 2 + 1) * dx
    if len(atm_cd) == 2:
        out = []
        for s in steps:
            d = np.zeros(eq.shape)
            d[atm_cd[0], atm_cd[1]] += s
            out.append(eq + d)
        return np.array(out)
    out = []
    for s1 in steps:
        for s2 in steps:
            d = np.zeros(eq.shape)
            d[atm_cd[0], atm_cd[1]] += s1
            d[atm_cd[2], atm_cd[3]] += s2
            out.append(eq + d)
    return np.array(out)


def hessian(eq, pot, dx=1e-3, pd=5, po=3):
    """harmonic_analysis.py:54-91: each element from its own potential call."""
    m = eq.shape[0]
    n = 3 * m
    h = np.zeros((n, n))
    flat = np.ravel_multi_index(np.array(np.triu_indices(n, 1)), [n, n])
    info = np.array(np.unravel_index(flat, (m, 3, m, 3))).T
    for k, ij in enumerate(itt.combinations(np.arange(n), 2)):
        v = np.reshape(pot(stencil(eq, info[k], dx, po)), (po, po))
        w = W1[po] / dx
        h[ij] = np.dot(w, np.dot(w, v))
    h = h + h.T
    for k, (a, c) in enumerate(itt.product(np.arange(m), np.arange(3))):
        h[k, k] = np.dot(W2[pd] / dx ** 2, pot(stencil(eq, (a, c), dx, pd)))
    return h


def gf_frequencies(m_h, m_o, re, theta, k_r, k_t):
    """Harmonic frequencies (hartree) of a bent [H, H, O] molecule with diagonal force
    constants k_r (each O-H stretch) and k_t (bend) by Wilson's GF method: eigenvalues of
    G F in the internal coordinates (r1, r2, theta) at the equilibrium geometry."""
    c, s = np.cos(theta), np.sin(theta)
    g11 = 1 / m_h + 1 / m_o
    g12 = c / m_o
    g13 = -s / (re * m_o)
    g33 = 2 / (re * re * m_h) + (2 - 2 * c) / (re * re * m_o)
    G = np.array([[g11, g12, g13], [g12, g11, g13], [g13, g13, g33]])
    F = np.diag([k_r, k_r, k_t])
    lam = np.sort(np.linalg.eigvals(G @ F).real)
    return np.sqrt(lam)
