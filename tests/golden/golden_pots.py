"""Potentials handed to the reference by ``gen_golden.py`` (as ``Potential_Direct``), kept as a
user potential module: the GPU tests load them through the same managers
(``Potential``, ``Potential_NoMP``, ``Potential_Direct``) a reference script uses.

``pes_ho_x``: V = 0.5 k x^2 on the first coordinate, k = m omega^2 (the 1-D oscillator of the
benchmark); ``pes_morse_bend``: two O-H Morse stretches and a harmonic bend of an [H, H, O]
walker.  Written in plain NumPy, independently of the oracle.
"""
import numpy as np

M_H, M_O, OMEGA = 1837.15, 29156.95, 0.013669
K_HO = M_H * OMEGA ** 2


def pes_ho_x(cds):
    x = cds[:, 0, 0]
    return 0.5 * K_HO * x ** 2


def pes_morse_bend(cds):
    De, a, re, kb, th0 = 0.1916, 1.2, 1.81, 0.16, np.deg2rad(104.5)
    u = cds[:, 0] - cds[:, 2]
    v = cds[:, 1] - cds[:, 2]
    r1 = np.sqrt(np.sum(u * u, axis=1))
    r2 = np.sqrt(np.sum(v * v, axis=1))
    c = np.clip(np.sum(u * v, axis=1) / (r1 * r2), -1.0, 1.0)
    th = np.arccos(c)
    return De * (1 - np.exp(-a * (r1 - re))) ** 2 + De * (1 - np.exp(-a * (r2 - re))) ** 2 \
        + 0.5 * kb * (th - th0) ** 2


def water_start(n, seed):
    t = np.deg2rad(104.5)
    eq = np.array([[1.81, 0.0, 0.0], [1.81 * np.cos(t), 1.81 * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
    return eq[None] + np.random.RandomState(seed).normal(0.0, 0.05, size=(n, 3, 3))


class CmModel:
    """A stand-in learned model for the ``NN_Potential`` convention (a ``predict`` on
    walkers that returns wavenumbers, as call_sample_model.py does): the Morse+bend surface
    in cm^-1."""

    def predict(self, cds):
        return pes_morse_bend(cds) / 4.556335281212229e-6


def nn_pes(cds, model):
    """potential_function(cds, model) of an NN_Potential: model output in cm^-1 -> hartree."""
    return model.predict(cds) * 4.556335281212229e-6
