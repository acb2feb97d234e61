"""Pins of the oracle's potentials, trial functions and derivatives by mathematics
(closed forms, central differences, limits), independent of the formulas under test."""
import numpy as np
import pytest

from oracle import pes as P
from oracle.eigen import discrete_tau_energy


def rand_water(n, seed=0, noise=0.08):
    return P.water_eq()[None] + np.random.RandomState(seed).normal(0, noise, size=(n, 3, 3))


def central_diff(f, x, h=1e-4):
    g = np.zeros_like(x)
    for a in range(x.shape[1]):
        for d in range(x.shape[2]):
            xp = x.copy()
            xm = x.copy()
            xp[:, a, d] += h
            xm[:, a, d] -= h
            g[:, a, d] = (f(xp) - f(xm)) / (2 * h)
    return g


def central_diff2(f, x, h=1e-4):
    s = np.zeros_like(x)
    f0 = f(x)
    for a in range(x.shape[1]):
        for d in range(x.shape[2]):
            xp = x.copy()
            xm = x.copy()
            xp[:, a, d] += h
            xm[:, a, d] -= h
            s[:, a, d] = (f(xp) - 2 * f0 + f(xm)) / h ** 2
    return s


def test_morse_bend_equilibrium_and_limits():
    eq = P.water_eq()[None]
    assert abs(P.morse_bend(eq)[0]) < 1e-15
    # stretch one bond far away: Morse term -> De
    far = eq.copy()
    far[0, 0, 0] = 60.0
    assert abs(P.morse_bend(far)[0] - P.MB_DE - 0.5 * P.MB_KB * (0.0) ** 2) < 1e-9 + 0.5 * P.MB_KB * 0.2
    # pure bend of 10 degrees, bonds at re: V = 0.5 kb (10 deg)^2
    t = P.MB_THETA0 + np.deg2rad(10)
    g = np.array([[[P.MB_RE, 0, 0], [P.MB_RE * np.cos(t), P.MB_RE * np.sin(t), 0], [0, 0, 0]]])
    assert np.isclose(P.morse_bend(g)[0], 0.5 * P.MB_KB * np.deg2rad(10) ** 2, rtol=1e-12)
    # rigid rotation and translation invariance
    x = rand_water(16)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    np.testing.assert_allclose(P.morse_bend(x @ R.T + 3.0), P.morse_bend(x), rtol=1e-11, atol=1e-14)


def test_morse_bend_harmonic_frequency():
    """omega_r = a sqrt(2 De / mu): the curvature of the Morse stretch at re."""
    p = P.water_trial_params()
    h = 1e-4
    v = lambda r: P.MB_DE * (1 - np.exp(-P.MB_A * (r - P.MB_RE))) ** 2
    k = (v(P.MB_RE + h) - 2 * v(P.MB_RE) + v(P.MB_RE - h)) / h ** 2
    assert np.isclose(np.sqrt(k / p['mu']), p['omega_r'], rtol=1e-6)
    assert np.isclose(p['omega_r'] / P.WAVENUMBERS, 3921.7, atol=0.5)
    assert np.isclose(p['omega_t'] / P.WAVENUMBERS, 1662.2, atol=0.5)


@pytest.mark.parametrize('fn,grad,x', [
    (P.morse_bend, P.morse_bend_grad, rand_water(20, 1)),
    (P.cluster6, P.cluster6_grad, P.cluster6_eq()[None] + np.random.RandomState(2).normal(0, 0.08, (20, 6, 3))),
])
def test_pes_gradients_vs_central_differences(fn, grad, x):
    np.testing.assert_allclose(grad(x), central_diff(fn, x), rtol=1e-5, atol=1e-8)


def test_bond_and_angle_derivatives_vs_central_differences():
    x = rand_water(12, 3)
    r, dr, d2r = P.bond(x, 0, 2)
    np.testing.assert_allclose(dr, central_diff(lambda c: P.bond(c, 0, 2)[0], x), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(d2r, central_diff2(lambda c: P.bond(c, 0, 2)[0], x), rtol=1e-4, atol=1e-6)
    th, dth, d2th = P.angle(x, 0, 2, 1)
    np.testing.assert_allclose(dth, central_diff(lambda c: P.angle(c, 0, 2, 1)[0], x), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(d2th, central_diff2(lambda c: P.angle(c, 0, 2, 1)[0], x), rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize('trial,x', [
    (P.water_trial(), rand_water(10, 4)),
    (P.cluster6_trial(), P.cluster6_eq()[None] + np.random.RandomState(5).normal(0, 0.06, (10, 6, 3))),
])
def test_gauss_product_derivatives_vs_central_differences(trial, x):
    d1, d2 = trial.derivs(x)
    psi = trial.trial(x)
    np.testing.assert_allclose(d1, central_diff(trial.trial, x) / psi[:, None, None], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(d2, central_diff2(trial.trial, x, h=2e-4) / psi[:, None, None], rtol=1e-3, atol=1e-2)
    np.testing.assert_allclose(np.log(psi), trial.log_trial(x), rtol=1e-12, atol=1e-12)


def test_harmonic_trial_closed_forms():
    t = P.HarmonicTrial1D()
    x = np.linspace(-0.3, 0.3, 21)[:, None, None]
    d1, d2 = t.derivs(x)
    np.testing.assert_allclose(d1[:, 0, 0], -t.alpha * x[:, 0, 0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d2[:, 0, 0], t.alpha ** 2 * x[:, 0, 0] ** 2 - t.alpha, rtol=1e-10, atol=1e-8)


def test_finite_diff_is_second_order_accurate():
    """imp_samp.py:55-76 central differences against the analytic HO derivatives."""
    t = P.HarmonicTrial1D()
    x = np.linspace(-0.2, 0.2, 9)[:, None, None].copy()
    f1, f2, psi = P.finite_diff(x.copy(), t.trial)
    d1, d2 = t.derivs(x)
    np.testing.assert_allclose(f1 / psi[:, None, None], d1, rtol=0, atol=1e-4 * t.alpha ** 2)
    np.testing.assert_allclose(f2 / psi[:, None, None], d2, rtol=0, atol=1e-4 * t.alpha ** 3)


def test_discrete_tau_energy_limits():
    k = P.M_H * P.HO_OMEGA ** 2
    e10 = discrete_tau_energy(P.M_H, k, 10.0) / P.WAVENUMBERS
    e01 = discrete_tau_energy(P.M_H, k, 0.1) / P.WAVENUMBERS
    assert abs(e01 - P.HO_OMEGA / 2 / P.WAVENUMBERS) < 1e-3
    assert abs(e10 - 1498.834) < 1e-2
    # Trotter error is O(dt^2)
    e5 = discrete_tau_energy(P.M_H, k, 5.0) / P.WAVENUMBERS
    w2 = P.HO_OMEGA / 2 / P.WAVENUMBERS
    assert np.isclose((w2 - e10) / (w2 - e5), 4.0, rtol=0.05)


def test_oracle_c0_zpe_statistics():
    """A short FP64 oracle run of the 1-D oscillator lands on the exact discrete-tau energy."""
    from oracle.dmc import LegacyDraws, OracleDMC
    k = P.M_H * P.HO_OMEGA ** 2
    o = OracleDMC(potential=lambda c: P.harmonic_x(c, k), masses=[P.M_H], start_structures=np.zeros((1, 1, 3)),
                  num_walkers=2000, num_timesteps=1500, delta_t=10, equil_steps=10 ** 9, wfn_every=10 ** 9,
                  draws=LegacyDraws(0))
    o.propagate()
    zpe = o.zpe_cm()
    assert abs(zpe - 1498.834) < 15.0


@pytest.mark.parametrize('which', ['water', 'cluster'])
def test_closed_form_weighted_laplacian_matches_per_component_derivatives(which):
    """The closed-form sum_a (1/m_a) lap_a ln psi used by the FP32 importance-sampled move
    equals the per-component sum of the derivatives the reference's trial functions return."""
    if which == 'water':
        tr, m = P.water_trial(), [P.M_H, P.M_H, P.M_O]
        x = P.water_eq()[None] + np.random.RandomState(3).normal(0, 0.08, (500, 3, 3))
    else:
        tr, m = P.cluster6_trial(), [P.M_O, P.M_H, P.M_H, P.M_O, P.M_H, P.M_H]
        x = P.cluster6_eq()[None] + np.random.RandomState(5).normal(0, 0.08, (500, 6, 3))
    d1, d2 = tr.derivs(x)
    im = np.repeat(1 / np.asarray(m), 3)
    n = len(x)
    ref = (im * d2.reshape(n, -1)).sum(1)
    fast = (im * d1.reshape(n, -1) ** 2).sum(1) + tr.weighted_laplacian(x, m)
    np.testing.assert_allclose(fast, ref, rtol=1e-11, atol=0)
