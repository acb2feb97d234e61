"""Exact ground-state energy of the reference's discrete-time-step scheme, 1-D oscillator.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

One reference step (pyvibdmc.py:540-547 then :391-431) maps the walker density rho to
exp(-dt*(V(x) - E)) * (K_dt * rho)(x), with K_dt the Gaussian of variance dt/m.  The
stationary population therefore grows as the largest eigenvalue lambda of
exp(-dt V/2) K_dt exp(-dt V/2) (same spectrum as exp(-dt V) K_dt), and the mean vref of
a long run converges to E = -ln(lambda)/dt.  Here lambda is computed on a uniform grid.
Pinned by the continuum limit dt -> 0 (E -> omega/2) in the tests.
"""
import numpy as np


def discrete_tau_energy(mass, k, dt, half_width=None, n=2001):
    omega = np.sqrt(k / mass)
    if half_width is None:
        half_width = 12.0 / np.sqrt(mass * omega)
    x = np.linspace(-half_width, half_width, n)
    h = x[1] - x[0]
    s2 = dt / mass
    kern = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * s2)) * (h / np.sqrt(2 * np.pi * s2))
    half = np.exp(-0.5 * dt * 0.5 * k * x * x)
    a = half[:, None] * kern * half[None, :]
    lam = np.linalg.eigvalsh(a)[-1]
    return -np.log(lam) / dt
