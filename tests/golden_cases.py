"""Shared helpers that rebuild a golden case's configuration for the oracle and for the GPU.

The golden files are written by ``tests/golden/gen_golden.py`` from the reference run.
"""
import os

import numpy as np

from oracle import pes as P
from oracle.dmc import LegacyDraws, OracleDMC

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden')

CASES = ['c0_discrete', 'c0_desc', 'c1_discrete', 'c1_continuous', 'c1_continuous_001',
         'c1_continuous_upper', 'is_ho1d', 'is_morse1d_dt10', 'is_water', 'is_water_fd']

# descendant-weighting schedule of each case (equil_steps, wfn_every, desc_wt_steps)
SCHED = {
    'c0_discrete': (10 ** 6, 10 ** 6, 1), 'c0_desc': (10, 40, 15),
    'c1_discrete': (30, 60, 25), 'c1_continuous': (20, 50, 20),
    'c1_continuous_001': (10 ** 6, 10 ** 6, 1), 'c1_continuous_upper': (10 ** 6, 10 ** 6, 1),
    'is_ho1d': (10 ** 6, 10 ** 6, 1), 'is_morse1d_dt10': (10 ** 6, 10 ** 6, 1),
    'is_water': (10 ** 6, 10 ** 6, 1), 'is_water_fd': (10 ** 6, 10 ** 6, 1),
}


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False))


def oh_mass():
    return P.reduced_mass_au(P.MASS_O_AMU, P.MASS_H_AMU)


def potential_for(g):
    kind = str(g['pes'])
    if kind == 'ho_x':
        k = float(g['k'])
        return lambda c: P.harmonic_x(c, k)
    if kind == 'morse_bend':
        return P.morse_bend
    if kind == 'oh_stretch_harm':
        # harmonicOscillator1D.py:65-69
        m, w = oh_mass(), 3700. * P.WAVENUMBERS
        return lambda c: P.ho_reference_style(c, m, w)
    if kind == 'oh_stretch_morse':
        # morse_osc_1d.py:4-12
        m = oh_mass()
        w = 3704.5 * P.WAVENUMBERS
        wx = 75.3 * P.WAVENUMBERS
        De = (w ** 2 / (4 * wx))
        a = np.sqrt(m * (w ** 2.) / 2. / De)
        return lambda c: np.squeeze(De * np.square((1 - np.exp(-a * c))))
    raise KeyError(kind)


class _FDTrial:
    """ImpSampManager_NoMP with no deriv_function: FD then divide by psi
    (imp_samp_manager.py:202-208)."""

    def __init__(self, f):
        self.f = f

    def trial(self, c):
        return self.f(c)

    def derivs(self, c):
        d1, d2, psi = P.finite_diff(c, self.f)
        return d1 / psi[:, None, None], d2 / psi[:, None, None]


class _FnTrial:
    def __init__(self, f, d):
        self.f, self.d = f, d

    def trial(self, c):
        return self.f(c)

    def derivs(self, c):
        return self.d(c)


def trial_for(g):
    if 'trial' not in g:
        return None
    kind = str(g['trial'])
    if kind == 'harm_trial_wfn':
        return P.HarmonicTrial1D()
    from tests.golden import golden_trials as gt
    if kind == 'water_gauss':
        return _FnTrial(gt.water_trial, gt.water_derivs)
    if kind == 'water_gauss_fd':
        return _FDTrial(gt.water_trial)
    raise KeyError(kind)


def oracle_for(name, g, draws=None, record_idx=True):
    eq, every, dw = SCHED[name]
    kw = {}
    if 'thresh_lower' in g:
        kw['weighting'] = 'continuous'
        kw['cont_wt_thresh'] = ([float(g['thresh_lower']), float(g['thresh_upper'])]
                                if 'thresh_upper' in g else float(g['thresh_lower']))
    tr = trial_for(g)
    return OracleDMC(potential=potential_for(g), masses=g['masses'], start_structures=g['start'],
                     num_walkers=int(g['num_walkers']), num_timesteps=int(g['num_timesteps']),
                     delta_t=float(g['delta_t']), equil_steps=eq, wfn_every=every, desc_wt_steps=dw,
                     trial=tr, imp_samp_oned=(str(g.get('trial', '')) == 'harm_trial_wfn'),
                     draws=draws or LegacyDraws(int(g['seed'])), record_idx=record_idx, **kw)


def split_counts(g):
    """Per-step count vectors from the flat golden array."""
    flat = g['counts_flat']
    out, o = [], 0
    for n in g['pop_in']:
        out.append(flat[o:o + n].astype(np.int64))
        o += n
    return out


def split_src(g):
    flat = g['src_flat']
    out, o = [], 0
    for n in g['pop_in']:
        out.append(flat[o:o + n] + np.arange(n))
        o += n
    return out
