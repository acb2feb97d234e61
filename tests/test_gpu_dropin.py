"""The reference's call surface on the GPU: user Python potentials and trial wave functions
through the reference's manager classes, branch_every > 1, run(), and the global NumPy
draw stream.

* Host trial functions (``ImpSampManager`` / ``ImpSampManager_NoMP`` over
  ``tests/golden/golden_trials.py``, with analytic and with finite-difference derivatives):
  the device proposes, the host evaluates psi and its derivatives, the device runs the
  Metropolis test, the caches and branching.  Against the trajectories the reference itself
  wrote (``is_water``, ``is_water_fd``): populations, births and deaths bit for bit, final
  coordinates bit for bit (the proposal is computed in the reference's operation order and
  the derivatives are the host's), vref within 1e-10 hartree, effective time within 1e-12.
* Host potentials (``Potential`` with a 2-process pool, ``Potential_NoMP``,
  ``Potential_Direct``, ``NN_Potential``): the split path against golden trajectories or the
  oracle.
* ``branch_every = 2`` (discrete, continuous, importance-sampled) against the oracle.
* ``run()``: final checkpoint, sim_info, ZPE; on failure a checkpoint and the reference's
  error.
* ``rng='numpy'``: the global legacy NumPy stream after ``np.random.seed``, as the reference
  draws it.
"""
import os

import numpy as np
import pytest

from oracle import pes as P
from oracle.dmc import LegacyDraws, OracleDMC
from tests.golden_cases import GOLDEN, load, oracle_for, split_counts

pytestmark = pytest.mark.gpu

M3 = [P.M_H, P.M_H, P.M_O]
FAR = dict(equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)


def golden_pots():
    import sys
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    import golden_pots
    return golden_pots


def check_against_golden(sim, g, exact_coords=True):
    T = int(g['num_timesteps'])
    np.testing.assert_allclose(sim._vref_vs_tau, g['vref'], rtol=0, atol=1e-10)
    b, d = sim.trajectory_counts(0, T)
    np.testing.assert_array_equal(b, g['births'])
    np.testing.assert_array_equal(d, g['deaths'])
    pops = np.array([c.sum() for c in split_counts(g)])
    np.testing.assert_array_equal(sim._pop_vs_tau.astype(np.int64), pops)
    coords, _ = sim._get_walkers()
    if exact_coords:
        np.testing.assert_array_equal(coords, g['final_coords'])
    else:
        np.testing.assert_allclose(coords, g['final_coords'], rtol=0, atol=1e-9)
    if 'eff_ts' in g:
        np.testing.assert_allclose(sim.eff_ts, g['eff_ts'], rtol=0, atol=1e-12)


def water_sim(g, potential, imp_samp=None, **kw):
    import pyvibdmc_amd as pv
    return pv.DMC_Sim(num_walkers=int(g['num_walkers']), num_timesteps=int(g['num_timesteps']),
                      atoms=['H', 'H', 'O'], masses=M3, delta_t=float(g['delta_t']), potential=potential,
                      start_structures=g['start'], imp_samp=imp_samp, precision='fp64',
                      rng=LegacyDraws(int(g['seed'])), write_files=False, **FAR, **kw)


@pytest.mark.parametrize('case,manager,pot', [
    ('is_water', 'NoMP', 'Direct'),
    ('is_water', 'MP', 'NoMP'),
    ('is_water_fd', 'NoMP', 'Pool'),
    ('is_water_fd', 'NoMP', 'device'),
])
def test_host_trial_importance_sampling_matches_reference(case, manager, pot):
    import pyvibdmc_amd as pv
    gp = golden_pots()
    g = load(case)
    deriv = 'water_derivs' if case == 'is_water' else None
    if pot == 'Direct':
        potential = pv.Potential_Direct(gp.pes_morse_bend)
    elif pot == 'NoMP':
        potential = pv.Potential_NoMP('pes_morse_bend', GOLDEN, 'golden_pots.py')
    elif pot == 'Pool':
        potential = pv.Potential('pes_morse_bend', GOLDEN, 'golden_pots.py', num_cores=2)
    else:
        potential = pv.MorseBendTriatomic()
    if manager == 'NoMP':
        imp = pv.ImpSampManager_NoMP('water_trial', GOLDEN, 'golden_trials.py', deriv_function=deriv)
    else:
        imp = pv.ImpSampManager('water_trial', GOLDEN, 'golden_trials.py', potential, deriv_function=deriv)
    sim = water_sim(g, potential, imp)
    sim.propagate()
    # the device potential differs from NumPy's arccos / exp in the last ulps
    check_against_golden(sim, g, exact_coords=(pot != 'device'))
    if pot == 'Pool':
        potential.mp_close()


@pytest.mark.parametrize('case,pot', [('c1_discrete', 'Direct'), ('c1_continuous', 'Pool'),
                                      ('c0_discrete', 'NoMP')])
def test_host_potential_matches_reference(case, pot):
    """The split path (device move, host potential on a host copy, device branching)."""
    import pyvibdmc_amd as pv
    from tests.golden_cases import SCHED
    gp = golden_pots()
    g = load(case)
    fn = 'pes_morse_bend' if str(g['pes']) == 'morse_bend' else 'pes_ho_x'
    if pot == 'Direct':
        potential = pv.Potential_Direct(getattr(gp, fn))
    elif pot == 'NoMP':
        potential = pv.Potential_NoMP(fn, GOLDEN, 'golden_pots.py')
    else:
        potential = pv.Potential(fn, GOLDEN, 'golden_pots.py', num_cores=2)
    eq, every, dw = SCHED[case]
    kw = {}
    if 'thresh_lower' in g:
        kw = dict(weighting='continuous', cont_wt_thresh=float(g['thresh_lower']), start_cont_wts=g['start_wts'])
    m = g['masses'].shape[0]
    sim = pv.DMC_Sim(num_walkers=int(g['num_walkers']), num_timesteps=int(g['num_timesteps']),
                     atoms=['H', 'H', 'O'] if m == 3 else ['X'], masses=list(g['masses']) if m == 3 else float(g['masses'][0]),
                     delta_t=float(g['delta_t']), potential=potential, start_structures=g['start'],
                     precision='fp64', rng=LegacyDraws(int(g['seed'])), write_files=False,
                     equil_steps=eq, wfn_every=every, desc_wt_steps=dw, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    np.testing.assert_allclose(sim._vref_vs_tau, g['vref'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(sim._pop_vs_tau, g['pop'], rtol=1e-12, atol=0)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, g['final_coords'])
    if wts is not None:
        np.testing.assert_allclose(wts, g['final_wts'], rtol=1e-12, atol=0)
    for i, (step, rec) in enumerate(getattr(sim, 'wfns', [])):
        assert step == int(g[f'wfn{i}_step'])
        np.testing.assert_array_equal(rec['coords'], g[f'wfn{i}_coords'])
    if pot == 'Pool':
        potential.mp_close()


def test_nn_potential_matches_oracle():
    """NN_Potential: the function receives the walkers and the model object
    (potential_manager.py:177-214); against the oracle driven by the same function."""
    import pyvibdmc_amd as pv
    gp = golden_pots()
    model = gp.CmModel()
    g = load('c1_discrete')
    ora = oracle_for('c1_discrete', g, record_idx=False)
    ora.potential = lambda c: gp.nn_pes(c, model)
    ora.propagate()
    pot = pv.NN_Potential('nn_pes', GOLDEN, 'golden_pots.py', model)
    sim = water_sim(g, pot)
    sim.propagate()
    np.testing.assert_array_equal(sim._pop_vs_tau, ora._pop_vs_tau)
    np.testing.assert_allclose(sim._vref_vs_tau, ora._vref_vs_tau, rtol=0, atol=1e-10)
    np.testing.assert_array_equal(sim._get_walkers()[0], ora.coords)


def start_water(n, seed):
    return P.water_eq()[None] + np.random.RandomState(seed).normal(0, 0.05, (n, 3, 3))


@pytest.mark.parametrize('mode', ['discrete', 'continuous', 'is_discrete', 'is_continuous_host'])
def test_branch_every_two_matches_oracle(mode):
    """branch_every = 2 (pyvibdmc.py:139, 828-837): branching on even steps only; continuous
    weights still multiply on the odd steps; with importance sampling the effective time
    advances on every continuous step but only on branching steps for discrete weighting."""
    import pyvibdmc_amd as pv
    from tests.golden_cases import _FnTrial
    gp = golden_pots()
    import golden_trials as gt
    n, T, seed = 400, 40, 21
    cont = 'continuous' in mode
    is_ = mode.startswith('is')
    kw = dict(weighting='continuous', cont_wt_thresh=0.2) if cont else {}
    ora = OracleDMC(potential=P.morse_bend, masses=M3, start_structures=start_water(n, 3), num_walkers=n,
                    num_timesteps=T, delta_t=5, branch_every=2, draws=LegacyDraws(seed),
                    trial=_FnTrial(gt.water_trial, gt.water_derivs) if is_ else None,
                    equil_steps=10 ** 6, wfn_every=10 ** 6, **kw)
    ora.propagate()
    if mode == 'is_discrete':
        trial = pv.water_trial()  # the device form of the same Gaussian product
        potential = pv.MorseBendTriatomic()
    elif mode == 'is_continuous_host':
        trial = pv.ImpSampManager_NoMP('water_trial', GOLDEN, 'golden_trials.py', deriv_function='water_derivs')
        potential = pv.Potential_Direct(gp.pes_morse_bend)
    else:
        trial = None
        potential = pv.MorseBendTriatomic()
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=T, atoms=['H', 'H', 'O'], masses=M3, delta_t=5,
                     potential=potential, start_structures=start_water(n, 3), imp_samp=trial, branch_every=2,
                     precision='fp64', rng=LegacyDraws(seed), write_files=False, **FAR, **kw)
    sim.propagate()
    np.testing.assert_allclose(sim._vref_vs_tau, ora._vref_vs_tau, rtol=0, atol=1e-10)
    np.testing.assert_allclose(sim._pop_vs_tau, ora._pop_vs_tau, rtol=1e-12, atol=0)
    c, w = sim._get_walkers()
    atol = 1e-9 if mode == 'is_discrete' else 0.0  # device PES / trial vs NumPy: last ulps
    np.testing.assert_allclose(c, ora.coords, rtol=0, atol=atol)
    if w is not None:
        np.testing.assert_allclose(w, ora.cont_wts, rtol=1e-12, atol=0)
    if is_:
        np.testing.assert_allclose(sim.eff_ts, ora.eff_ts, rtol=0, atol=1e-12)
        if not cont:
            assert np.all(sim.eff_ts[1::2] == 0.0)


def test_numpy_rng_follows_the_reference_stream():
    """rng='numpy' draws from the global legacy stream after np.random.seed, as the reference
    does (pyvibdmc.py:392, 544-546): the golden c1_discrete trajectory bit for bit."""
    import pyvibdmc_amd as pv
    g = load('c1_discrete')
    from tests.golden_cases import SCHED
    eq, every, dw = SCHED['c1_discrete']
    np.random.seed(int(g['seed']))
    sim = pv.DMC_Sim(num_walkers=int(g['num_walkers']), num_timesteps=int(g['num_timesteps']),
                     atoms=['H', 'H', 'O'], masses=M3, delta_t=float(g['delta_t']),
                     potential=pv.MorseBendTriatomic(), start_structures=g['start'], precision='fp64',
                     rng='numpy', write_files=False, equil_steps=eq, wfn_every=every, desc_wt_steps=dw,
                     chkpt_every=10 ** 9)
    sim.propagate()
    np.testing.assert_allclose(sim._vref_vs_tau, g['vref'], rtol=0, atol=1e-10)
    np.testing.assert_array_equal(sim._get_walkers()[0], g['final_coords'])


def test_run_writes_results_and_checkpoints(tmp_path, capsys):
    """run() (pyvibdmc.py:878-932): a final checkpoint, the sim_info record with the time axis
    and the ZPE printout; a restart continues from the checkpoint."""
    import pyvibdmc_amd as pv
    k = P.M_H * P.HO_OMEGA ** 2
    kw = dict(sim_name='ho', output_folder=str(tmp_path), num_walkers=2000, num_timesteps=300, equil_steps=100,
              wfn_every=100, desc_wt_steps=20, chkpt_every=100, atoms=['X'], masses=P.M_H, delta_t=10,
              start_structures=np.zeros((1, 1, 3)), seed=4)
    sim = pv.DMC_Sim(potential=pv.HarmonicOscillator1D(k=k), **kw)
    sim.run()
    out = capsys.readouterr().out
    assert 'Approximate ZPE' in out
    zpe = float(out.split('Approximate ZPE')[1].split()[0])
    assert abs(zpe - np.average(sim._vref_vs_tau[75:] / P.WAVENUMBERS)) < 1e-9
    assert os.path.exists(tmp_path / 'chkpts' / 'ho_299.pickle')
    info = sim.sim_info
    assert info['vref_vs_tau'].shape == (300, 2)
    np.testing.assert_array_equal(info['vref_vs_tau'][:, 0], np.arange(300) * 10.0)
    assert sorted(os.listdir(tmp_path / 'wfns')) == ['ho_wfn_100ts.hdf5', 'ho_wfn_200ts.hdf5']
    # the reference's reader sees what run() saved (pyvibdmc.py:861-869, 918-922)
    si = pv.SimInfo(str(tmp_path / 'ho_sim_info.hdf5'))
    np.testing.assert_array_equal(si.get_vref(), info['vref_vs_tau'])
    np.testing.assert_array_equal(si.get_pop(), info['pop_vs_tau'])
    np.testing.assert_array_equal(si.get_atom_masses(), [P.M_H])
    for (step, rec), t in zip(sim.wfns, (100, 200)):
        cds, dw = si.get_wfns(t)
        assert step == t
        np.testing.assert_array_equal(cds, rec['coords'])
        np.testing.assert_array_equal(dw, rec['desc_wts'])
    re = pv.dmc_restart(pv.HarmonicOscillator1D(k=k), str(tmp_path), 'ho', additional_timesteps=50)
    re.propagate()
    assert re._vref_vs_tau.shape == (350,) and np.all(re._vref_vs_tau[300:] != 0.0)


def test_run_failure_checkpoints_and_raises(tmp_path, capsys):
    """A massive death inside run(): the final checkpoint is written and the reference's
    ValueError propagates (pyvibdmc.py:892-932, 409-413)."""
    import pyvibdmc_amd as pv
    sim = pv.DMC_Sim(sim_name='boom', output_folder=str(tmp_path), num_walkers=1000, num_timesteps=50,
                     atoms=['X'], masses=P.M_H, delta_t=10, potential=pv.HarmonicOscillator1D(k=1.0),
                     start_structures=np.full((1, 1, 3), 3.0), DEBUG_alpha=1e6, seed=1, equil_steps=10 ** 9,
                     wfn_every=10 ** 9, chkpt_every=10 ** 9)
    with pytest.raises(ValueError, match='Massive walker birth or death event'):
        sim.run()
    out = capsys.readouterr().out
    assert 'Dumping a final checkpoint' in out
    assert any(f.startswith('boom_') for f in os.listdir(tmp_path / 'chkpts'))
