"""GPU parity: the HIP step kernels, fed the reference's own draws, reproduce the
reference's trajectories (golden vectors written by the reference, tests/golden/).

Bar: bit-exact populations per step, births and deaths per step (discrete) or number of
splits per step (continuous), final coordinates (plain moves only add the injected
displacement, so they are bit-exact whenever every branching decision matched); vref within
1e-10 hartree.  With importance sampling the GPU's analytic drift differs from the
reference's in the last ulps, so coordinates are compared to 1e-9 bohr.
"""
import numpy as np
import pytest

from tests.golden_cases import load, split_counts

pytestmark = pytest.mark.gpu

GPU_CASES = ['c0_discrete', 'c0_desc', 'c1_discrete', 'c1_continuous', 'c1_continuous_001', 'is_ho1d',
             'is_morse1d_dt10', 'is_water']


@pytest.mark.parametrize('name', GPU_CASES)
def test_gpu_matches_reference(name):
    from tests.gpu_cases import gpu_sim_for
    g = load(name)
    sim = gpu_sim_for(name, g)
    sim.propagate()
    T = int(g['num_timesteps'])
    np.testing.assert_allclose(sim._vref_vs_tau, g['vref'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(sim._pop_vs_tau, g['pop'], rtol=1e-12, atol=0)
    b, d = sim.trajectory_counts(0, T)
    if 'counts_flat' in g:
        np.testing.assert_array_equal(b, g['births'])
        np.testing.assert_array_equal(d, g['deaths'])
        pops = np.array([c.sum() for c in split_counts(g)])
        np.testing.assert_array_equal(sim._pop_vs_tau.astype(np.int64), pops)
    else:
        np.testing.assert_array_equal(b, g['n_branched'])
    coords, wts = sim._get_walkers()
    assert coords.shape == g['final_coords'].shape
    if 'trial' in g:
        np.testing.assert_allclose(coords, g['final_coords'], rtol=0, atol=1e-9)
        np.testing.assert_allclose(sim.eff_ts, g['eff_ts'], rtol=0, atol=1e-12)
    else:
        np.testing.assert_array_equal(coords, g['final_coords'])
    if wts is not None:
        np.testing.assert_allclose(wts, g['final_wts'], rtol=1e-12, atol=0)
    wf = getattr(sim, 'wfns', [])
    assert len(wf) == int(g['n_wfn'])
    for i, (step, rec) in enumerate(wf):
        assert step == int(g[f'wfn{i}_step'])
        np.testing.assert_array_equal(rec['coords'], g[f'wfn{i}_coords'])
        np.testing.assert_allclose(rec['desc_wts'], g[f'wfn{i}_desc_wts'], rtol=1e-12, atol=1e-13)
        if 'parent_wts' in rec:
            # continuous weights are products of exp() factors: ulp-level differences
            np.testing.assert_allclose(rec['parent_wts'], g[f'wfn{i}_parent_wts'], rtol=1e-12, atol=0)


def test_gpu_fp32_tracks_fp64():
    """FP32 coordinates / FP64 reductions on the same injected draws stay on the reference
    trajectory for a short run (vref to 1e-6 hartree)."""
    from tests.gpu_cases import gpu_sim_for
    g = load('c1_discrete')
    sim = gpu_sim_for('c1_discrete', g, precision='fp32')
    sim.propagate()
    T = 40
    np.testing.assert_allclose(sim._vref_vs_tau[:T], g['vref'][:T], rtol=0, atol=1e-6)


def test_large_population_matches_oracle():
    """1.5M walkers, FP64, the reference's draws injected: 2 930 tiles of 512 walkers, so the
    compaction's look-back crosses several 512-tile windows.  Two steps match the oracle bit
    for bit in populations, births, deaths and coordinates (np.repeat order)."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n = 1_500_000
    masses = [1837.15, 1837.15, 29156.95]
    start = P.water_eq()[None] + np.random.RandomState(21).normal(0, 0.05, (n, 3, 3))
    orc = OracleDMC(potential=P.morse_bend, masses=masses, start_structures=start, num_walkers=n,
                    num_timesteps=2, delta_t=5, equil_steps=10 ** 9, wfn_every=10 ** 9, draws=LegacyDraws(5))
    orc.propagate()
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=2, atoms=['H', 'H', 'O'], masses=masses, delta_t=5,
                     potential=pv.MorseBendTriatomic(), start_structures=start, precision='fp64',
                     rng=LegacyDraws(5), write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9,
                     chkpt_every=10 ** 9)
    sim.propagate()
    np.testing.assert_array_equal(sim._pop_vs_tau, orc._pop_vs_tau)
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    b, d = sim.trajectory_counts(0, 2)
    np.testing.assert_array_equal(b, orc.rec['births'])
    np.testing.assert_array_equal(d, orc.rec['deaths'])
    coords, _ = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)


@pytest.mark.parametrize('n,steps', [(100_000, 120), (400_000, 2)])
def test_continuous_large_population_matches_oracle(n, steps):
    """Continuous weighting at n walkers, FP64, the reference's draws injected, start weights
    log-uniform over [e^-4.8, e^2] so that about 3 % of the walkers fall below the 0.01
    threshold in the first step (12 000 splits at 400k), a few hundred a step after that.
    The first step has no candidate list and runs the interval replay; the later ones take
    the lists the weight kernel builds.  Splits per step, vref, weights and coordinates match
    the oracle's sequential argmax/halve/clone loop (pyvibdmc.py:340-356)."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    masses = [1837.15, 1837.15, 29156.95]
    rs = np.random.RandomState(22)
    start = P.water_eq()[None] + rs.normal(0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-4.8, 2.0, n))
    kw = dict(masses=masses, start_structures=start, num_walkers=n, num_timesteps=steps, delta_t=5,
              equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01,
              start_cont_wts=w0)
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(8), **kw)
    orc.propagate()
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(8), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    assert nb[0] > 0.02 * n and how[0] == 2  # no list yet in the first step
    assert np.isin(how[1:], [1, 5, 6]).all(), how
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)


def test_continuous_replay_with_more_kills_than_the_lds_sort_matches_oracle():
    """600k walkers, FP64, the reference's draws injected; start weights log-uniform over
    [e^-2, e^2] (scaled to mean 1) with 20 % in a band just above the 0.01 threshold
    (log-uniform over a tenth of an e-fold), so that 5 000 - 14 000 walkers cross it per step
    -- more than the LDS replay serves (1.25 K > 4096).  The first step runs the interval replay (no list yet); every later step
    with more than 3 277 kills takes the global-memory candidate replay; splits per step,
    vref, weights and coordinates match the oracle's sequential loop."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 600_000, 5
    masses = [1837.15, 1837.15, 29156.95]
    rs = np.random.RandomState(23)
    start = P.water_eq()[None] + rs.normal(0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-2.0, 2.0, n))  # the heavy end spread over octaves
    band = rs.rand(n) < 0.2
    w0[band] = 0.01 * np.exp(rs.uniform(0.0, 0.1, int(band.sum())))
    w0[~band] *= (n - w0[band].sum()) / w0[~band].sum()  # mean weight 1: vref starts near v-bar
    kw = dict(masses=masses, start_structures=start, num_walkers=n, num_timesteps=steps, delta_t=5,
              equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01,
              start_cont_wts=w0)
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(9), **kw)
    orc.propagate()
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(9), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    print('splits', nb, 'replay', how, flush=True)
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    big = nb + nb // 4 > 4096
    assert how[0] == 2 and big.sum() >= 2, (nb, how)
    assert np.isin(how[1:][big[1:]], [7, 8]).all() and np.isin(how[1:][~big[1:]], [1, 5, 6]).all(), (nb, how)
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)


def test_continuous_replay_from_a_million_walker_steady_state_matches_oracle():
    """1M walkers, FP64: 300 Philox steps of continuous weighting (threshold 0.01) reach the
    steady state, about 16 000 kills per step.  There the split pairs of each step pile up
    just above its cut, so the candidates crowd a narrow mantissa range.  From those walkers and
    weights, four steps with the reference's draws injected: the first runs the interval
    replay (no list yet), the later ones the global-memory candidate replay; splits per step,
    vref, weights and coordinates match the oracle's sequential loop."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 1_000_000, 4
    masses = [1837.15, 1837.15, 29156.95]
    start = P.water_eq()[None] + np.random.RandomState(5).normal(0, 0.05, (n, 3, 3))
    common = dict(masses=masses, num_walkers=n, delta_t=5, equil_steps=10 ** 9, wfn_every=10 ** 9,
                  weighting='continuous', cont_wt_thresh=0.01)
    warm = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64', seed=3,
                      start_structures=start, num_timesteps=300, write_files=False, chkpt_every=10 ** 9, **common)
    warm.propagate()
    nb_w, how_w = warm.trajectory_counts(0, 300)
    print('warm-up splits', nb_w[-5:], 'replay', np.unique(how_w, return_counts=True), flush=True)
    assert nb_w[-50:].min() > 10_000, nb_w[-50:]
    c0, w0 = warm._get_walkers()
    del warm
    kw = dict(start_structures=c0, start_cont_wts=w0, num_timesteps=steps, **common)
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(13), **kw)
    orc.propagate()
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(13), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    print('splits', nb, 'replay', how, flush=True)
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    assert how[0] == 2 and np.isin(how[1:], [7, 8]).all(), (nb, how)
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)


@pytest.mark.timeout(300)
def test_continuous_importance_sampled_replay_at_scale_matches_oracle():
    """600k walkers, FP64, importance sampling with the water Gaussian-product trial, the
    reference's draws injected; half the start weights in a band a hundredth of an e-fold
    above the threshold, so that 20 000 - 34 000 walkers cross it per step (the local energy
    keeps the weight factors near 1): the interval replay in the first step, the
    global-memory candidate replay after.  Splits per step, vref and effective time steps
    match the oracle; coordinates to 1e-9 bohr and weights to 1e-10 (the analytic drift
    differs from the reference's in the last ulps, as in the golden IS cases)."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 600_000, 3
    masses = [1837.15, 1837.15, 29156.95]
    rs = np.random.RandomState(29)
    start = P.water_eq()[None] + rs.normal(0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-2.0, 2.0, n))
    band = rs.rand(n) < 0.5
    w0[band] = 0.01 * np.exp(rs.uniform(0.0, 0.01, int(band.sum())))
    w0[~band] *= (n - w0[band].sum()) / w0[~band].sum()
    kw = dict(masses=masses, start_structures=start, num_walkers=n, num_timesteps=steps, delta_t=5,
              equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01,
              start_cont_wts=w0)
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), imp_samp=pv.water_trial(),
                     precision='fp64', rng=LegacyDraws(31), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    print('splits', nb, 'replay', how, flush=True)
    assert how[0] == 2 and np.isin(how[1:], [7, 8]).all(), (nb, how)
    coords, wts = sim._get_walkers()
    vref, eff = sim._vref_vs_tau.copy(), np.asarray(sim.eff_ts).copy()
    del sim
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(31), trial=P.water_trial(), **kw)
    orc.propagate()
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    np.testing.assert_allclose(vref, orc._vref_vs_tau, rtol=0, atol=1e-10)
    np.testing.assert_allclose(eff, orc.eff_ts, rtol=0, atol=1e-12)
    np.testing.assert_allclose(coords, orc.coords, rtol=0, atol=1e-9)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-10, atol=0)


@pytest.mark.timeout(600)
def test_continuous_replay_beyond_32768_kills_per_step_matches_oracle():
    """2M walkers, FP64, the reference's draws injected, the spread start weights of the
    600k case (log-uniform over [e^-2, e^2], 20 % in a band just above the threshold): about
    48 000 kills in the first step (interval replay) and 38 000 in the second (global-memory
    candidate replay), both above the 32 768 a step was limited to before; splits per step,
    vref, weights and coordinates match the oracle's sequential loop."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 2_000_000, 2
    masses = [1837.15, 1837.15, 29156.95]
    rs = np.random.RandomState(23)
    start = P.water_eq()[None] + rs.normal(0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-2.0, 2.0, n))
    band = rs.rand(n) < 0.2
    w0[band] = 0.01 * np.exp(rs.uniform(0.0, 0.1, int(band.sum())))
    w0[~band] *= (n - w0[band].sum()) / w0[~band].sum()
    kw = dict(masses=masses, start_structures=start, num_walkers=n, num_timesteps=steps, delta_t=5,
              equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01,
              start_cont_wts=w0)
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(9), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    print('splits', nb, 'replay', how, flush=True)
    assert nb.min() > 32768 and how[0] == 2 and how[1] in (7, 8), (nb, how)
    coords, wts = sim._get_walkers()
    vref = sim._vref_vs_tau.copy()
    del sim
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(9), **kw)
    orc.propagate()
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    np.testing.assert_allclose(vref, orc._vref_vs_tau, rtol=0, atol=1e-10)
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)


def test_continuous_replay_in_the_first_steps_of_a_million_walker_run_matches_oracle():
    """1M walkers at the equilibrium geometry + N(0, 0.05 bohr), weights 1, FP64, the
    reference's draws injected, 14 steps: the kills grow from none to a few thousand per step
    while the weights are still bunched, so the candidate lists (threshold from the last
    step's cut, or list B below it) hold tens of thousands of walkers.  The LDS form declines
    them as too long and the global-memory form serves those steps (code 8) although the
    kills fit the LDS sort; splits per step, vref, weights and coordinates match the oracle."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 1_000_000, 14
    masses = [1837.15, 1837.15, 29156.95]
    start = P.water_eq()[None] + np.random.RandomState(0).normal(0.0, 0.05, (n, 3, 3))
    kw = dict(masses=masses, start_structures=start, num_walkers=n, num_timesteps=steps, delta_t=5,
              equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01)
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(21), **kw)
    orc.propagate()
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(21), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    print('splits', nb, 'replay', how, flush=True)
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    small = (nb > 0) & (nb + nb // 4 <= 4096)
    assert (small & (how == 8)).any(), (nb, how)
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)


def test_continuous_replay_with_a_cut_inside_the_list_matches_oracle():
    """A tenth of the walkers start in one dense cluster of near-equal heavy weights (5.0,
    spread 1e-6) and 3 % below the threshold.  After the first step (interval replay) the
    candidate list holds more walkers than the replay sorts, so the replay cuts inside it
    (path 5) before the later steps take the whole list.  Everything matches the oracle's
    sequential split loop."""
    import pyvibdmc_amd as pv
    from oracle import pes as P
    from oracle.dmc import LegacyDraws, OracleDMC
    n, steps = 100_000, 8
    rs = np.random.RandomState(30)
    start = P.water_eq()[None] + rs.normal(0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-3.0, 0.5, n))
    sel = rs.rand(n)
    heavy = sel < 0.10
    w0[heavy] = 5.0 * (1.0 + 1e-6 * rs.normal(size=heavy.sum()))
    w0[(sel >= 0.10) & (sel < 0.13)] = 0.005
    kw = dict(masses=[1837.15, 1837.15, 29156.95], start_structures=start, num_walkers=n, num_timesteps=steps,
              delta_t=5, equil_steps=10 ** 9, wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01,
              start_cont_wts=w0)
    orc = OracleDMC(potential=P.morse_bend, draws=LegacyDraws(9), **kw)
    orc.propagate()
    sim = pv.DMC_Sim(atoms=['H', 'H', 'O'], potential=pv.MorseBendTriatomic(), precision='fp64',
                     rng=LegacyDraws(9), write_files=False, chkpt_every=10 ** 9, **kw)
    sim.propagate()
    nb, how = sim.trajectory_counts(0, steps)
    np.testing.assert_array_equal(nb, orc.rec['n_branched'])
    assert how[0] == 2 and 5 in how.tolist(), how
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    coords, wts = sim._get_walkers()
    np.testing.assert_array_equal(coords, orc.coords)
    np.testing.assert_allclose(wts, orc.cont_wts, rtol=1e-12, atol=0)
