"""CPU baseline: the oracle driven with the reference's process-pool potential manager.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``); used as the ``cpu_baseline`` leg of
``bench.py``.

Structure of the reference's CPU path: diffusion, branching and vref run serially in NumPy
in the parent process (pyvibdmc.py:540-547, 391-431, 651-661); only the potential is
farmed out, with ``np.array_split(cds, num_cores)`` -> ``Pool.map`` -> ``np.concatenate``
(simulation_utilities/potential_manager.py:83-88).
"""
import multiprocessing as mp
import os
import time

import numpy as np

from . import pes as P
from .dmc import OracleDMC


class PoolPotential:
    """``Potential.getpot`` of the reference: split, map over a fork pool, concatenate."""

    def __init__(self, fn, num_cores):
        self.fn = fn
        self.num_cores = num_cores
        self.pool = mp.get_context('fork').Pool(num_cores) if num_cores > 1 else None

    def __call__(self, cds):
        if self.pool is None:
            return self.fn(cds)
        return np.concatenate(self.pool.map(self.fn, np.array_split(cds, self.num_cores)))

    def close(self):
        if self.pool is not None:
            self.pool.close()
            self.pool.join()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


class _Draws:
    def __init__(self, seed):
        self.rs = np.random.RandomState(seed)

    def displacement(self, sigmas, shape):
        n, m, d = shape
        return self.rs.normal(0.0, sigmas, size=(n, d, m)).transpose(0, 2, 1)

    def uniform(self, n):
        return self.rs.random_sample(n)


def time_plain_triatomic(num_walkers=1_000_000, steps=10, warmup=1, num_cores=None, seed=0, dt=5.0):
    """Walker-steps/s of the reference-structured FP64 CPU path on the triatomic Morse+bend
    PES, discrete weighting.  Returns (walker_steps_per_s, seconds, walker_steps, cores)."""
    cores = host_cores() if num_cores is None else num_cores
    pot = PoolPotential(P.morse_bend, cores)
    try:
        start = P.water_eq()[None] + np.random.RandomState(seed).normal(0.0, 0.05, size=(num_walkers, 3, 3))
        o = OracleDMC(potential=pot, masses=[P.M_H, P.M_H, P.M_O], start_structures=start,
                      num_walkers=num_walkers, num_timesteps=warmup + steps, delta_t=dt,
                      equil_steps=10 ** 9, wfn_every=10 ** 9, draws=_Draws(seed + 1))
        o.propagate(warmup)
        n0 = len(o.rec['pop_in'])
        t0 = time.perf_counter()
        o.propagate(steps)
        sec = time.perf_counter() - t0
        ws = int(np.sum(o.rec['pop_in'][n0:]))
    finally:
        pot.close()
    return ws / sec, sec, ws, cores
