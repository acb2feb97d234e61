"""Trial wave functions handed to the reference's ImpSampManager_NoMP by ``gen_golden.py``.

A product of Gaussians in the two O-H stretches and the H-O-H bend of a [H, H, O]
walker, written independently of the oracle, with analytic first and diagonal second
Cartesian derivatives divided by psi (the ``deriv_function`` contract,
reference ``docs/imp_samp.rst:26-29``).
"""
import numpy as np

M_H, M_O = 1837.15, 29156.95
DE, A, RE, KB, TH0 = 0.1916, 1.2, 1.81, 0.16, np.deg2rad(104.5)
_mu = M_H * M_O / (M_H + M_O)
ALPHA_R = _mu * A * np.sqrt(2.0 * DE / _mu)
_g = 2.0 / (M_H * RE ** 2) + (2.0 / RE ** 2 - 2.0 * np.cos(TH0) / RE ** 2) / M_O
ALPHA_T = np.sqrt(KB * _g) / _g


def _internals(cds):
    h1, h2, o = cds[:, 0], cds[:, 1], cds[:, 2]
    u = h1 - o
    v = h2 - o
    r1 = np.sqrt(np.sum(u * u, axis=1))
    r2 = np.sqrt(np.sum(v * v, axis=1))
    c = np.clip(np.sum(u * v, axis=1) / (r1 * r2), -1.0, 1.0)
    return u, v, r1, r2, c


def water_trial(cds):
    _, _, r1, r2, c = _internals(cds)
    th = np.arccos(c)
    q = ALPHA_R * (r1 - RE) ** 2 + ALPHA_R * (r2 - RE) ** 2 + ALPHA_T * (th - TH0) ** 2
    return np.exp(-0.5 * q)


def water_derivs(cds):
    """(dpsi/psi, d2psi/psi), each (N,3,3)."""
    u, v, r1, r2, c = _internals(cds)
    n = len(cds)
    eu = u / r1[:, None]
    ev = v / r2[:, None]
    s = np.sqrt(1.0 - c * c)
    th = np.arccos(c)
    g = np.zeros((n, 3, 3))
    h = np.zeros((n, 3, 3))
    # stretches: d r = +-e, d2 r = (1-e^2)/r
    for e, r, a in ((eu, r1, 0), (ev, r2, 1)):
        x = (r - RE)[:, None]
        d2 = (1.0 - e * e) / r[:, None]
        g[:, a] += -ALPHA_R * x * e
        g[:, 2] += ALPHA_R * x * e
        h[:, a] += -ALPHA_R * (e * e + x * d2)
        h[:, 2] += -ALPHA_R * (e * e + x * d2)
    # bend through c = cos(theta)
    C = c[:, None]
    dcu = (ev - C * eu) / r1[:, None]
    dcv = (eu - C * ev) / r2[:, None]
    duu = (-2.0 * eu * ev - C + 3.0 * C * eu * eu) / (r1 * r1)[:, None]
    dvv = (-2.0 * eu * ev - C + 3.0 * C * ev * ev) / (r2 * r2)[:, None]
    duv = (1.0 - eu * eu - ev * ev + C * eu * ev) / (r1 * r2)[:, None]
    S = s[:, None]
    x = (th - TH0)[:, None]
    for dc, d2c, a in ((dcu, duu, 0), (dcv, dvv, 1), (-(dcu + dcv), duu + 2.0 * duv + dvv, 2)):
        dth = -dc / S
        d2th = -d2c / S - C * dc * dc / (S * S * S)
        g[:, a] += -ALPHA_T * x * dth
        h[:, a] += -ALPHA_T * (dth * dth + x * d2th)
    return g, g * g + h
