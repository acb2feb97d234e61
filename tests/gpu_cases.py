"""Builds the GPU-side DMC_Sim for a golden case (device potential + device trial)."""
import numpy as np

import pyvibdmc_amd as pv
from oracle.dmc import LegacyDraws
from tests.golden_cases import SCHED


def device_potential(g):
    kind = str(g['pes'])
    if kind == 'ho_x':
        return pv.HarmonicOscillator1D(k=float(g['k']))
    if kind == 'morse_bend':
        return pv.MorseBendTriatomic()
    if kind == 'oh_stretch_harm':
        return pv.HarmonicOscillator1D.oh_stretch()
    if kind == 'oh_stretch_morse':
        return pv.MorseOscillator1D.oh_stretch()
    raise KeyError(kind)


def device_trial(g):
    if 'trial' not in g:
        return None
    kind = str(g['trial'])
    if kind == 'harm_trial_wfn':
        return pv.HarmonicTrial1D()
    if kind == 'water_gauss':
        from tests.golden import golden_trials as gt
        return pv.GaussianProductTrial(3, bonds=[(0, 2, gt.ALPHA_R, gt.RE), (1, 2, gt.ALPHA_R, gt.RE)],
                                       angles=[(0, 2, 1, gt.ALPHA_T, gt.TH0)])
    raise KeyError(kind)


def gpu_sim_for(name, g, precision='fp64', rng=None):
    eq, every, dw = SCHED[name]
    kw = {}
    if 'thresh_lower' in g:
        kw['weighting'] = 'continuous'
        kw['cont_wt_thresh'] = float(g['thresh_lower'])
        kw['start_cont_wts'] = g['start_wts']
    masses = g['masses']
    m = masses.shape[0]
    atoms = ['X'] * m
    if m == 1 and str(g.get('pes')) in ('oh_stretch_harm', 'oh_stretch_morse'):
        atoms = ['O-H']
        mass_arg = None
    elif m == 1:
        mass_arg = float(masses[0])
    else:
        atoms = ['H', 'H', 'O']
        mass_arg = list(masses)
    return pv.DMC_Sim(sim_name=name, num_walkers=int(g['num_walkers']), num_timesteps=int(g['num_timesteps']),
                      equil_steps=eq, wfn_every=every, desc_wt_steps=dw, chkpt_every=10 ** 9,
                      atoms=atoms, masses=mass_arg, delta_t=float(g['delta_t']), potential=device_potential(g),
                      start_structures=g['start'], imp_samp=device_trial(g),
                      imp_samp_oned=(str(g.get('trial', '')) == 'harm_trial_wfn'),
                      precision=precision, rng=rng if rng is not None else LegacyDraws(int(g['seed'])),
                      write_files=False, **kw)
