"""Write golden DMC trajectories by running the reference pyvibdmc in this container.

Usage (development container only; the reference does not exist on GPU hosts):

    python tests/golden/gen_golden.py [--ref /root/reference]

What it does
------------
* Inserts a stand-in ``h5py`` module that keeps ``SimArchivist.save_h5`` outputs in
  memory (the reference imports h5py at module load, ``simulation_utilities/sim_archive.py:1``;
  h5py is not installed here).
* Imports ``pyvibdmc`` from the reference tree without writing bytecode.
* Runs short seeded simulations through the reference's own ``DMC_Sim.run()`` with a
  subclass that only *records*: before each ``birth_or_death`` it points the reference's
  own descendant tracker at ``arange(N)`` so that afterwards it holds the survivor
  indices (discrete, ``pyvibdmc.py:417,429-430``) or the resolved clone sources
  (continuous, ``pyvibdmc.py:349-350``), then restores the real tracker.  No draw is added
  or removed, so the trajectory is the reference's own.
* Saves inputs and outputs as ``tests/golden/<case>.npz``.  Only these data files are
  committed; no reference source or bytecode enters the repository.

The draws themselves are not stored: tests regenerate them with
``numpy.random.RandomState(seed)`` in the reference's order.
"""
import argparse
import os
import shutil
import sys
import tempfile
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SAVED = {}


class _H5File:
    def __init__(self, fname, mode='r'):
        self.fname = fname
        self.data = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        SAVED[os.path.basename(self.fname)] = self.data
        return False

    def create_dataset(self, key, data):
        self.data[key] = np.array(data)


def import_reference(ref):
    h5 = types.ModuleType('h5py')
    h5.File = _H5File
    sys.modules['h5py'] = h5
    sys.dont_write_bytecode = True
    os.environ.setdefault('MPLBACKEND', 'Agg')
    sys.path.insert(0, ref)
    import pyvibdmc as pv
    return pv


# ---- configuration PESs (written here, independent of the oracle) ----------------------
from golden_pots import M_H, M_O, OMEGA, K_HO, pes_ho_x, pes_morse_bend, water_start  # noqa: E402


def make_recorder(pv):
    class Recorder(pv.DMC_Sim):
        def birth_or_death(self):
            if not hasattr(self, 'g_rec'):
                self.g_rec = dict(pop_in=[], counts=[], src=[], dt_factor=[], births=[], deaths=[],
                                  nbr=[])
            n_in = len(self._walker_coords)
            self.g_rec['pop_in'].append(n_in)
            if self.impsamp_manager is not None:
                self.g_rec['dt_factor'].append(self.dt_factor)
            real_on, real_wf = self._desc_wt, self._who_from
            self._desc_wt, self._who_from = True, np.arange(n_in)
            try:
                out = super().birth_or_death()
                got = self._who_from
            finally:
                self._desc_wt = real_on
                self._who_from = real_wf
            if real_on:
                self._who_from = real_wf[got]
            if self.weighting == 'discrete':
                self.g_rec['counts'].append(np.bincount(got, minlength=n_in).astype(np.uint8))
                self.g_rec['births'].append(out[0])
                self.g_rec['deaths'].append(out[1])
            else:
                self.g_rec['src'].append(got.astype(np.int32))
                self.g_rec['nbr'].append(out[0])
            return out
    # the reference pickles the simulation object at the end of run(); make the class findable
    Recorder.__qualname__ = 'Recorder'
    Recorder.__module__ = __name__
    globals()['Recorder'] = Recorder
    return Recorder


def run_case(pv, name, outdir, seed, sim_kwargs, extra):
    Recorder = make_recorder(pv)
    SAVED.clear()
    work = tempfile.mkdtemp(prefix='golden_')
    try:
        np.random.seed(seed)
        sim = Recorder(sim_name=name, output_folder=os.path.join(work, 'out'), chkpt_every=10 ** 9,
                       **sim_kwargs)
        start = np.array(sim._walker_coords, copy=True)
        start_w = None if sim._cont_wts is None else np.array(sim._cont_wts, copy=True)
        sim.run()
    finally:
        shutil.rmtree(work, ignore_errors=True)
    r = sim.g_rec
    T = sim.num_timesteps
    data = dict(seed=seed, start=start, vref=sim._vref_vs_tau.copy(), pop=sim._pop_vs_tau.copy(),
                pop_in=np.array(r['pop_in'], dtype=np.int32), final_coords=np.asarray(sim._walker_coords),
                final_pots=np.asarray(sim._walker_pots), masses=np.asarray(sim.masses, dtype=np.float64),
                num_walkers=sim.num_walkers, num_timesteps=T, delta_t=float(sim.delta_t))
    if start_w is not None:
        data['start_wts'] = start_w
        data['final_wts'] = np.asarray(sim._cont_wts)
        data['src_flat'] = np.concatenate(r['src']) - np.concatenate([np.arange(len(s)) for s in r['src']])
        data['n_branched'] = np.array(r['nbr'], dtype=np.int32)
    else:
        data['counts_flat'] = np.concatenate(r['counts'])
        data['births'] = np.array(r['births'], dtype=np.int64)
        data['deaths'] = np.array(r['deaths'], dtype=np.int64)
    if r['dt_factor']:
        data['dt_factor'] = np.array(r['dt_factor'])
        data['eff_ts'] = np.asarray(sim.eff_ts)
    wf = sorted((k for k in SAVED if '_wfn_' in k), key=lambda k: int(k.split('_wfn_')[1].split('ts')[0]))
    for i, k in enumerate(wf):
        step = int(k.split('_wfn_')[1].split('ts')[0])
        data[f'wfn{i}_step'] = step
        for key, val in SAVED[k].items():
            data[f'wfn{i}_{key}'] = val
    data['n_wfn'] = len(wf)
    data.update(extra)
    path = os.path.join(outdir, f'{name}.npz')
    np.savez_compressed(path, **data)
    print(f'wrote {path}: T={T}, final N={len(sim._walker_coords)}, vref[-1]={sim._vref_vs_tau[-1]:.10f}')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--ref', default='/root/reference')
    ap.add_argument('--out', default=HERE)
    args = ap.parse_args()
    pv = import_reference(args.ref)
    pypots = os.path.join(args.ref, 'pyvibdmc', 'sample_potentials', 'PythonPots')

    # C0-mini: 1-D oscillator on x of an (N,1,3) walker, discrete, dt = 10.
    run_case(pv, 'c0_discrete', args.out, 0, dict(
        weighting='discrete', num_walkers=256, num_timesteps=200, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['X'], masses=M_H, delta_t=10,
        potential=pv.Potential_Direct(pes_ho_x), start_structures=np.zeros((1, 1, 3)), log_every=10 ** 6),
        dict(pes='ho_x', k=K_HO))

    # discrete with descendant-weighting windows.
    run_case(pv, 'c0_desc', args.out, 1, dict(
        weighting='discrete', num_walkers=256, num_timesteps=120, equil_steps=10, wfn_every=40,
        desc_wt_steps=15, atoms=['X'], masses=M_H, delta_t=10,
        potential=pv.Potential_Direct(pes_ho_x), start_structures=np.zeros((1, 1, 3)), log_every=10 ** 6),
        dict(pes='ho_x', k=K_HO))

    # C1-mini: triatomic Morse+bend, continuous weights, stress threshold, desc windows.
    run_case(pv, 'c1_continuous', args.out, 2, dict(
        weighting='continuous', num_walkers=512, num_timesteps=150, equil_steps=20, wfn_every=50,
        desc_wt_steps=20, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), start_structures=water_start(512, 1234),
        cont_wt_thresh=0.6, log_every=10 ** 6),
        dict(pes='morse_bend', thresh_lower=0.6))

    # continuous with the configuration threshold 0.01.
    run_case(pv, 'c1_continuous_001', args.out, 3, dict(
        weighting='continuous', num_walkers=512, num_timesteps=300, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), start_structures=water_start(512, 99),
        cont_wt_thresh=0.01, log_every=10 ** 6),
        dict(pes='morse_bend', thresh_lower=0.01))

    # continuous with lower and upper thresholds (reference test_run_dmc_cont pattern).
    run_case(pv, 'c1_continuous_upper', args.out, 4, dict(
        weighting='continuous', num_walkers=256, num_timesteps=100, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), start_structures=water_start(256, 7),
        cont_wt_thresh=[0.3, 2.0], log_every=10 ** 6),
        dict(pes='morse_bend', thresh_lower=0.3, thresh_upper=2.0))

    # discrete triatomic (plain headline kind, small).
    run_case(pv, 'c1_discrete', args.out, 5, dict(
        weighting='discrete', num_walkers=512, num_timesteps=150, equil_steps=30, wfn_every=60,
        desc_wt_steps=25, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), start_structures=water_start(512, 5),
        log_every=10 ** 6),
        dict(pes='morse_bend'))

    # IS on the reference's own 1-D oscillator and trial function (tests/test_imp_samp.py:9-49).
    ho = pv.Potential_NoMP(potential_function='oh_stretch_harm', potential_directory=pypots,
                           python_file='harmonicOscillator1D.py')
    imp = pv.ImpSampManager_NoMP(trial_function='trial_harm', trial_directory=pypots,
                                 python_file='harm_trial_wfn.py', deriv_function='derivative', chdir=True)
    run_case(pv, 'is_ho1d', args.out, 6, dict(
        weighting='discrete', num_walkers=300, num_timesteps=100, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['O-H'], delta_t=1, potential=ho, imp_samp=imp, imp_samp_oned=True,
        start_structures=np.zeros((1, 1, 1)), log_every=10 ** 6),
        dict(pes='oh_stretch_harm', trial='harm_trial_wfn'))

    # IS, 1-D Morse with dt = 10 (tests/test_imp_samp.py:89-124).
    mo = pv.Potential_NoMP(potential_function='oh_stretch_morse', potential_directory=pypots,
                           python_file='morse_osc_1d.py')
    run_case(pv, 'is_morse1d_dt10', args.out, 7, dict(
        weighting='discrete', num_walkers=300, num_timesteps=100, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['O-H'], delta_t=10, potential=mo, imp_samp=imp, imp_samp_oned=True,
        start_structures=np.zeros((1, 1, 1)), log_every=10 ** 6),
        dict(pes='oh_stretch_morse', trial='harm_trial_wfn'))

    # IS on the triatomic with analytic Gaussian-product trial derivatives (C2-mini).
    wimp = pv.ImpSampManager_NoMP(trial_function='water_trial', trial_directory=HERE,
                                  python_file='golden_trials.py', deriv_function='water_derivs')
    run_case(pv, 'is_water', args.out, 8, dict(
        weighting='discrete', num_walkers=256, num_timesteps=100, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), imp_samp=wimp,
        start_structures=water_start(256, 11), log_every=10 ** 6),
        dict(pes='morse_bend', trial='water_gauss'))

    # IS on the triatomic with the reference's finite-difference derivatives.
    fimp = pv.ImpSampManager_NoMP(trial_function='water_trial', trial_directory=HERE,
                                  python_file='golden_trials.py')
    run_case(pv, 'is_water_fd', args.out, 9, dict(
        weighting='discrete', num_walkers=64, num_timesteps=30, equil_steps=10 ** 6, wfn_every=10 ** 6,
        desc_wt_steps=1, atoms=['H', 'H', 'O'], masses=[M_H, M_H, M_O], delta_t=5,
        potential=pv.Potential_Direct(pes_morse_bend), imp_samp=fimp,
        start_structures=water_start(64, 13), log_every=10 ** 6),
        dict(pes='morse_bend', trial='water_gauss_fd'))


if __name__ == '__main__':
    main()
