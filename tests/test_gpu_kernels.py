"""Stand-alone device potentials and trial functions against the oracle, their hand-written
backward kernels against oracle gradients and FP64 central differences, and statistical
checks of the production (Philox) path."""
import numpy as np
import pytest

from oracle import pes as P

pytestmark = pytest.mark.gpu


def water(n, seed, noise=0.08):
    return P.water_eq()[None] + np.random.RandomState(seed).normal(0, noise, (n, 3, 3))


def cluster(n, seed, noise=0.08):
    return P.cluster6_eq()[None] + np.random.RandomState(seed).normal(0, noise, (n, 6, 3))


def cases():
    import pyvibdmc_amd as pv
    k = P.M_H * P.HO_OMEGA ** 2
    x1 = np.random.RandomState(1).normal(0, 0.2, (500, 1, 3))
    return [
        ('ho_x', pv.HarmonicOscillator1D(k=k), lambda c: P.harmonic_x(c, k), None, x1),
        ('morse_bend', pv.MorseBendTriatomic(), P.morse_bend, P.morse_bend_grad, water(2000, 2)),
        ('cluster6', pv.TwoMonomerCluster(), P.cluster6, P.cluster6_grad, cluster(2000, 3)),
        ('ps_water', pv.PartridgeSchwenkeWater(), P.ps_water, None,
         P.ps_eq()[None] + np.random.RandomState(4).normal(0, 0.1, (2000, 3, 3))),
    ]


@pytest.mark.parametrize('i', range(4))
def test_device_pes_matches_oracle_fp64_and_fp32(i):
    import torch
    name, dev, ref, _, x = cases()[i]
    v = dev.getpot(x)
    np.testing.assert_allclose(v, ref(x), rtol=1e-12, atol=1e-15)
    x32 = torch.from_numpy(x.astype(np.float32)).cuda()
    from pyvibdmc_amd import _native as N
    v32 = dev.eval_device(x32, precision=N.PVD_FP32).cpu().numpy()
    np.testing.assert_allclose(v32, ref(x.astype(np.float32).astype(np.float64)), rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize('i', [1, 2])
def test_pes_backward_kernel_matches_gradients(i):
    """PotentialEnergy autograd: backward kernel vs oracle analytic gradient and FP64 central differences."""
    import torch
    from pyvibdmc_amd.impsamp import PotentialEnergy
    name, dev, ref, grad, x = cases()[i]
    xt = torch.from_numpy(x[:200]).cuda().requires_grad_(True)
    v = PotentialEnergy.apply(xt, dev)
    v.sum().backward()
    g = xt.grad.cpu().numpy()
    np.testing.assert_allclose(g, grad(x[:200]), rtol=1e-9, atol=1e-12)
    h = 1e-4
    fd = np.zeros_like(x[:20])
    for a in range(x.shape[1]):
        for d in range(3):
            xp = x[:20].copy()
            xm = x[:20].copy()
            xp[:, a, d] += h
            xm[:, a, d] -= h
            fd[:, a, d] = (dev.getpot(xp) - dev.getpot(xm)) / (2 * h)
    np.testing.assert_allclose(g[:20], fd, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(dev.forces(x[:20]), -g[:20], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize('which', ['water', 'cluster', 'ho1d'])
def test_trial_kernels_match_oracle(which):
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.impsamp import TrialLogPsi
    if which == 'water':
        dev, ref, x = pv.water_trial(), P.water_trial(), water(1000, 4)
        dev.desc.c0 = 0.0
    elif which == 'cluster':
        dev, ref, x = pv.cluster6_trial(), P.cluster6_trial(), cluster(1000, 5)
    else:
        dev, ref, x = pv.HarmonicTrial1D(), P.HarmonicTrial1D(), np.random.RandomState(6).normal(0, 0.1, (500, 1, 1))
    d1, d2 = dev.call_derivs(x)
    r1, r2 = ref.derivs(x)
    np.testing.assert_allclose(d1, r1, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(d2, r2, rtol=1e-9, atol=1e-8)
    if which != 'ho1d':
        # ln psi forward / backward kernel against the oracle and central differences of ln psi
        xt = torch.from_numpy(x[:100]).cuda().requires_grad_(True)
        lp = TrialLogPsi.apply(xt, dev)
        np.testing.assert_allclose(lp.detach().cpu().numpy(), ref.log_trial(x[:100]) - ref.log_trial(x[:1])[0]
                                   + lp.detach().cpu().numpy()[0], rtol=1e-10, atol=1e-10)
        lp.sum().backward()
        np.testing.assert_allclose(xt.grad.cpu().numpy(), r1[:100], rtol=1e-10, atol=1e-10)
        h = 1e-4
        fd = np.zeros_like(x[:10])
        for a in range(x.shape[1]):
            for d in range(3):
                xp = x[:10].copy()
                xm = x[:10].copy()
                xp[:, a, d] += h
                xm[:, a, d] -= h
                fd[:, a, d] = (ref.log_trial(xp) - ref.log_trial(xm)) / (2 * h)
        np.testing.assert_allclose(r1[:10], fd, rtol=1e-6, atol=1e-7)


def test_c0_philox_zpe_matches_discrete_tau_energy():
    """Production path, 1-D oscillator: 5000 walkers, dt = 10, 2000 steps, FP64, 32 seeds.
    The mean over seeds of the ZPE estimate (mean vref over the last 3/4) matches the exact
    discrete-tau ground state 1498.83 cm^-1 within 3 standard errors, the standard error being
    at most 0.5 cm^-1.  The seeds are fixed, so the outcome is deterministic."""
    import pyvibdmc_amd as pv
    from oracle.eigen import discrete_tau_energy
    k = P.M_H * P.HO_OMEGA ** 2
    exact = discrete_tau_energy(P.M_H, k, 10.0) / P.WAVENUMBERS
    zs = []
    for seed in range(32):
        sim = pv.DMC_Sim(num_walkers=5000, num_timesteps=2000, atoms=['X'], masses=P.M_H, delta_t=10,
                         potential=pv.HarmonicOscillator1D(k=k), start_structures=np.zeros((1, 1, 3)),
                         precision='fp64', seed=seed, write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9,
                         chkpt_every=10 ** 9)
        sim.propagate()
        zs.append(sim.zpe())
        pop = sim._pop_vs_tau[500:]
        assert 4000 < pop.mean() < 5300
    zs = np.array(zs)
    err = zs.std(ddof=1) / np.sqrt(len(zs))
    assert err <= 0.5, err
    assert abs(zs.mean() - exact) < 3 * err, (zs.mean(), err, exact)


def test_massive_death_raises_reference_error():
    """A potential far above vref kills every walker: the reference's ValueError."""
    import pyvibdmc_amd as pv
    sim = pv.DMC_Sim(num_walkers=1000, num_timesteps=50, atoms=['X'], masses=P.M_H, delta_t=10,
                     potential=pv.HarmonicOscillator1D(k=1.0), start_structures=np.full((1, 1, 3), 3.0),
                     precision='fp64', seed=0, write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9,
                     chkpt_every=10 ** 9, DEBUG_alpha=1e6)
    with pytest.raises(ValueError, match='Massive walker birth or death event'):
        sim.propagate()


def test_restart_from_checkpoint_continues_the_same_draws(tmp_path):
    """A restart from the step-100 checkpoint reproduces the uninterrupted run: Philox draws
    are addressed by (walker slot, step), and the checkpoint keeps slot order."""
    import pyvibdmc_amd as pv
    kw = dict(num_walkers=2000, atoms=['H', 'H', 'O'], masses=[1837.15, 1837.15, 29156.95], delta_t=5,
              start_structures=water(2000, 7, 0.03), precision='fp32', seed=42, equil_steps=10 ** 9,
              wfn_every=10 ** 9, output_folder=str(tmp_path))
    a = pv.DMC_Sim(sim_name='a', num_timesteps=200, chkpt_every=100, potential=pv.MorseBendTriatomic(), **kw)
    a.propagate()                      # leaves the step-100 checkpoint on disk
    b = pv.dmc_restart(pv.MorseBendTriatomic(), str(tmp_path), 'a', additional_timesteps=100)
    assert b.cur_timestep == 100
    b.propagate()
    c = pv.DMC_Sim(sim_name='c', num_timesteps=300, chkpt_every=10 ** 9, potential=pv.MorseBendTriatomic(), **kw)
    c.propagate()
    np.testing.assert_array_equal(b._pop_vs_tau, c._pop_vs_tau)
    np.testing.assert_allclose(b._vref_vs_tau, c._vref_vs_tau, rtol=0, atol=1e-12)


def test_continuous_restart_from_checkpoint_continues_the_same_run(tmp_path):
    """Continuous weighting (20 000 walkers, FP64, threshold 0.01): a restart from the step-100
    checkpoint reproduces the uninterrupted run: same walkers, weights and vref.  The restarted
    run's first step has no candidate list yet and takes the interval replay, which is exact
    as well; the sums around the split may round differently there, hence vref to 1e-12 and
    the weights to 1e-12 relative."""
    import pyvibdmc_amd as pv
    n = 20000
    kw = dict(num_walkers=n, atoms=['H', 'H', 'O'], masses=[1837.15, 1837.15, 29156.95], delta_t=5,
              start_structures=water(n, 8, 0.05), precision='fp64', seed=43, equil_steps=10 ** 9,
              wfn_every=10 ** 9, weighting='continuous', cont_wt_thresh=0.01, output_folder=str(tmp_path))
    a = pv.DMC_Sim(sim_name='a', num_timesteps=200, chkpt_every=100, potential=pv.MorseBendTriatomic(), **kw)
    a.propagate()
    b = pv.dmc_restart(pv.MorseBendTriatomic(), str(tmp_path), 'a', additional_timesteps=100)
    assert b.cur_timestep == 100
    b.propagate()
    c = pv.DMC_Sim(sim_name='c', num_timesteps=300, chkpt_every=10 ** 9, potential=pv.MorseBendTriatomic(), **kw)
    c.propagate()
    nb_b, _ = b.trajectory_counts(100, 300)
    nb_c, _ = c.trajectory_counts(100, 300)
    assert nb_c.sum() > 1000
    np.testing.assert_array_equal(nb_b, nb_c)
    np.testing.assert_allclose(b._vref_vs_tau, c._vref_vs_tau, rtol=0, atol=1e-12)
    cb, wb = b._get_walkers()
    cc, wc = c._get_walkers()
    np.testing.assert_array_equal(cb, cc)
    np.testing.assert_allclose(wb, wc, rtol=1e-12, atol=0)


@pytest.mark.parametrize('mode', ['discrete', 'continuous', 'is'])
def test_graph_replay_matches_eager_launches(monkeypatch, mode):
    """pvd_sim_run replays a 64-step HIP graph; 300 steps (4 replays + remainder) must give
    the same populations, vref and walkers bit for bit as one launch at a time."""
    import pyvibdmc_amd as pv
    runs = []
    for graphs in ('1', '0'):
        monkeypatch.setenv('PVD_GRAPHS', graphs)
        kw = {}
        if mode == 'continuous':
            kw = dict(weighting='continuous', cont_wt_thresh=0.05)
        if mode == 'is':
            kw = dict(imp_samp=pv.water_trial())
        sim = pv.DMC_Sim(num_walkers=20000, num_timesteps=300, atoms=['H', 'H', 'O'],
                         masses=[1837.15, 1837.15, 29156.95], delta_t=5, potential=pv.MorseBendTriatomic(),
                         start_structures=water(20000, 12, 0.05), precision='fp32', seed=7, write_files=False,
                         equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9, **kw)
        sim.propagate()
        c, w = sim._get_walkers()
        runs.append((sim._pop_vs_tau.copy(), sim._vref_vs_tau.copy(), c, w))
    (p1, v1, c1, w1), (p2, v2, c2, w2) = runs
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(c1, c2)
    if w1 is not None:
        np.testing.assert_array_equal(w1, w2)


@pytest.mark.timeout(300)
def test_plain_fp32_production_step_tracks_fp64_at_a_million_walkers():
    """The headline step as it runs in production: water (triatomic Morse+bend), 1M walkers,
    dt = 5, discrete branching, Philox draws, the compact FP32 Box-Muller layout and the FP32
    branching weight (k_discrete<float,9>), 6 000 steps, against the FP64 run (FP64 draws and
    weights).  Same ZPE within 3 combined blocking errors, both within 2 cm^-1 of this surface's
    dt = 5 value (4676.5 cm^-1), and the population within 2 % of N0 throughout."""
    import pyvibdmc_amd as pv
    WN = 4.556335281212229e-6
    n, T = 1_000_000, 6000
    start = P.water_eq()[None] + np.random.RandomState(12).normal(0.0, 0.05, (n, 3, 3))
    out = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=n, num_timesteps=T, atoms=['H', 'H', 'O'], masses=[P.M_H, P.M_H, P.M_O],
                         delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start, equil_steps=10 ** 9,
                         wfn_every=10 ** 9, chkpt_every=10 ** 9, precision=prec, seed=12, write_files=False)
        sim.propagate()
        pop = np.asarray(sim._pop_vs_tau[:T], float)
        assert np.abs(pop / n - 1.0).max() < 0.02, (prec, pop.min(), pop.max())
        v = sim._vref_vs_tau[T // 4:] / WN
        se, b = 0.0, 1
        while len(v) // b >= 16:
            m = len(v) // b
            se = max(se, v[:m * b].reshape(m, b).mean(axis=1).std(ddof=1) / np.sqrt(m))
            b *= 2
        out[prec] = (v.mean(), se)
        del sim
    (z32, s32), (z64, s64) = out['fp32'], out['fp64']
    assert abs(z32 - z64) < 3 * np.hypot(s32, s64), out
    assert abs(z64 - 4676.5) < 2.0 and abs(z32 - 4676.5) < 2.0, out


def test_is_fp32_fast_trial_path_matches_fp64_zpe():
    """The FP32 importance-sampled move on the [H, H, O] trial uses gradients and a
    closed-form mass-weighted Laplacian; FP64 uses the per-component derivatives of the
    reference.  Independent Philox runs (200k walkers, 2000 steps) give the same ZPE within
    their statistical errors (a wrong Laplacian term shifts it by hundreds of cm^-1)."""
    import pyvibdmc_amd as pv
    WN = 4.556335281212229e-6
    z = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=200_000, num_timesteps=2000, atoms=['H', 'H', 'O'],
                         masses=[1837.15, 1837.15, 29156.95], delta_t=5, potential=pv.MorseBendTriatomic(),
                         start_structures=water(200_000, 31, 0.05), imp_samp=pv.water_trial(), precision=prec,
                         seed=8, write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
        sim.propagate()
        z[prec] = np.mean(sim._vref_vs_tau[500:]) / WN
    assert abs(z['fp32'] - z['fp64']) < 2.0, z
    assert 4670 < z['fp64'] < 4690, z


def test_c1_continuous_fp32_with_descendant_weighting_tracks_fp64():
    """The continuous-weighting workload: triatomic surface, 100k walkers, dt = 5, 10 000 steps,
    lower weight threshold 0.01, descendant weighting every 500 steps over 50 steps, FP32
    propagation against the FP64 run.  Both give the same ZPE within 3 combined blocking
    errors, the same number of descendant windows, and the same descendant-weighted mean O-H
    distance (the psi^2 expectation the windows exist for) within 0.002 bohr."""
    import pyvibdmc_amd as pv
    WN = 4.556335281212229e-6
    n, T = 100_000, 10_000
    start = P.water_eq()[None] + np.random.RandomState(0).normal(0.0, 0.05, (n, 3, 3))
    out = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=n, num_timesteps=T, atoms=['H', 'H', 'O'], masses=[P.M_H, P.M_H, P.M_O],
                         delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start,
                         weighting='continuous', cont_wt_thresh=0.01, equil_steps=500, wfn_every=500,
                         desc_wt_steps=50, chkpt_every=10 ** 9, precision=prec, seed=4, write_files=False)
        sim.propagate()
        v = sim._vref_vs_tau[T // 4:] / WN
        se = 0.0
        b = 1
        while len(v) // b >= 16:
            m = len(v) // b
            se = max(se, v[:m * b].reshape(m, b).mean(axis=1).std(ddof=1) / np.sqrt(m))
            b *= 2
        roh = []
        for _, rec in sim.wfns:
            c, w = rec['coords'], rec['desc_wts']
            assert np.all(w >= 0) and w.sum() > 0.5 * n
            r = 0.5 * (np.linalg.norm(c[:, 0] - c[:, 2], axis=1) + np.linalg.norm(c[:, 1] - c[:, 2], axis=1))
            roh.append(np.sum(w * r) / np.sum(w))
        out[prec] = (v.mean(), se, len(sim.wfns), np.mean(roh))
    (z32, s32, k32, r32), (z64, s64, k64, r64) = out['fp32'], out['fp64']
    assert k32 == k64 >= 18, out
    assert abs(z32 - z64) < 3 * np.hypot(s32, s64), out
    assert abs(r32 - r64) < 2e-3 and 1.80 < r64 < 1.90, out


@pytest.mark.timeout(300)
def test_continuous_fp32_at_a_million_walkers_tracks_fp64():
    """Continuous weighting at 1M walkers (about 16 000 kills per step, the global-memory
    candidate replay on nearly every step), dt = 5, 8 000 steps: the FP32 propagation and the
    FP64 run give the same ZPE within 3 combined blocking errors, and both sit on the plain
    dt = 5 value of this surface (4676.5 cm^-1, FP64 discrete at 1M walkers) within 2 cm^-1."""
    import pyvibdmc_amd as pv
    WN = 4.556335281212229e-6
    n, T = 1_000_000, 8000
    start = P.water_eq()[None] + np.random.RandomState(9).normal(0.0, 0.05, (n, 3, 3))
    out = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=n, num_timesteps=T, atoms=['H', 'H', 'O'], masses=[P.M_H, P.M_H, P.M_O],
                         delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start,
                         weighting='continuous', cont_wt_thresh=0.01, equil_steps=10 ** 9, wfn_every=10 ** 9,
                         chkpt_every=10 ** 9, precision=prec, seed=9, write_files=False)
        sim.propagate()
        nb, how = sim.trajectory_counts(0, T)
        assert np.isin(how[T // 2:], [7, 8]).mean() > 0.99, np.unique(how, return_counts=True)
        v = sim._vref_vs_tau[T // 4:] / WN
        se, b = 0.0, 1
        while len(v) // b >= 16:
            m = len(v) // b
            se = max(se, v[:m * b].reshape(m, b).mean(axis=1).std(ddof=1) / np.sqrt(m))
            b *= 2
        out[prec] = (v.mean(), se)
        del sim
    (z32, s32), (z64, s64) = out['fp32'], out['fp64']
    assert abs(z32 - z64) < 3 * np.hypot(s32, s64), out
    assert abs(z64 - 4676.5) < 2.0 and abs(z32 - 4676.5) < 2.0, out


def test_continuous_step_beyond_the_kill_limit_stops_the_run():
    """1.5M walkers, half of them with weights in a band just above the 0.01 threshold: about
    90 000 walkers cross it in the first step, more than the 65 536 a step may split.  The
    replay sets the status without touching the walkers, every later kernel of the step returns
    on it, and the run stops with a DMCError naming the split, at step 0."""
    import pyvibdmc_amd as pv
    from pyvibdmc_amd._native import DMCError
    n = 1_500_000
    rs = np.random.RandomState(41)
    start = P.water_eq()[None] + rs.normal(0.0, 0.05, (n, 3, 3))
    w0 = np.exp(rs.uniform(-2.0, 2.0, n))
    band = rs.rand(n) < 0.5
    w0[band] = 0.01 * np.exp(rs.uniform(0.0, 0.1, int(band.sum())))
    w0[~band] *= (n - w0[band].sum()) / w0[~band].sum()
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=5, atoms=['H', 'H', 'O'], masses=[P.M_H, P.M_H, P.M_O],
                     delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start, start_cont_wts=w0,
                     weighting='continuous', cont_wt_thresh=0.01, equil_steps=10 ** 9, wfn_every=10 ** 9,
                     chkpt_every=10 ** 9, precision='fp32', seed=3, write_files=False)
    with pytest.raises(DMCError, match='continuous split outside'):
        sim.propagate()
    assert sim.cur_timestep == 0


def test_is_fp32_cluster_matches_fp64_zpe():
    """Six-atom cluster with the per-monomer trial: the FP32 move (gradients + closed-form
    weighted Laplacian, runtime bond/angle lists) and the FP64 move (per-component
    derivatives) give the same ZPE within statistics (100k walkers, 1500 steps each)."""
    import pyvibdmc_amd as pv
    WN = 4.556335281212229e-6
    z = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=100_000, num_timesteps=1500, atoms=['O', 'H', 'H', 'O', 'H', 'H'],
                         masses=[P.M_O, P.M_H, P.M_H, P.M_O, P.M_H, P.M_H], delta_t=5,
                         potential=pv.TwoMonomerCluster(), start_structures=cluster(100_000, 41, 0.05),
                         imp_samp=pv.cluster6_trial(), precision=prec, seed=9, write_files=False,
                         equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
        sim.propagate()
        z[prec] = np.mean(sim._vref_vs_tau[500:]) / WN
    assert abs(z['fp32'] - z['fp64']) < 4.0, z


@pytest.mark.parametrize('atoms', [['H', 'D', 'O'], ['D', 'H', 'O']])
def test_is_fp32_isotopologue_tracks_fp64(atoms):
    """HOD / DHO with the [H, H, O] Gaussian trial: the single-precision move weights the
    drift and the Laplacian with each atom's own 1/m.  On the same injected draws the FP32
    run follows the FP64 run (vref to 2e-5 hartree for 30 steps); weighting the deuteron
    with the proton's 1/m moves the local energy by ~1e-3 hartree."""
    import pyvibdmc_amd as pv
    from oracle.dmc import LegacyDraws
    vref = {}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=2000, num_timesteps=30, atoms=atoms, delta_t=5,
                         potential=pv.MorseBendTriatomic(), start_structures=water(2000, 17, 0.05),
                         imp_samp=pv.water_trial(), precision=prec, rng=LegacyDraws(23), write_files=False,
                         equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
        sim.propagate()
        vref[prec] = sim._vref_vs_tau.copy()
    np.testing.assert_allclose(vref['fp32'], vref['fp64'], rtol=0, atol=2e-5)


@pytest.mark.parametrize('which', ['water', 'cluster'])
def test_harmonic_analysis_with_a_device_potential(which):
    """The Hessian's stencils evaluated by the device potential in one launch give the
    frequencies of the host NumPy potential (FP64 ulp differences, amplified by 1/dx^2)."""
    import pyvibdmc_amd as pv
    if which == 'water':
        eq, atoms, m = P.water_eq(), ['H', 'H', 'O'], [P.M_H, P.M_H, P.M_O]
        dev, host = pv.MorseBendTriatomic(), P.morse_bend
    else:
        eq, atoms, m = P.cluster6_eq(), ['O', 'H', 'H', 'O', 'H', 'H'], [P.M_O, P.M_H, P.M_H, P.M_O, P.M_H, P.M_H]
        dev, host = pv.TwoMonomerCluster(), P.cluster6
    fd, vd = pv.HarmonicAnalysis(eq, atoms, dev, masses=m).run()
    fh, vh = pv.HarmonicAnalysis(eq, atoms, pv.Potential_Direct(host), masses=m).run()
    np.testing.assert_allclose(fd[6:], fh[6:], rtol=0, atol=1e-3)



def test_ps_water_matches_the_reference_fortran():
    """The device Partridge-Schwenke surface against energies the reference's own Fortran
    wrote (tests/golden/ps_water.npz): FP64 to 1e-12 relative; FP32 to 2e-6 hartree (0.4 cm^-1)
    up to 0.5 hartree above the minimum.  Walker order [O, H, H] through the atom indices
    gives the same energies."""
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import _native as N
    from tests.golden_cases import GOLDEN
    g = np.load(GOLDEN + '/ps_water.npz')
    x, v = g['cds'], g['v']
    dev = pv.PartridgeSchwenkeWater()
    np.testing.assert_allclose(dev.getpot(x), v, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(pv.PartridgeSchwenkeWater(atoms=(1, 2, 0)).getpot(x[:, [2, 0, 1]]), v,
                               rtol=1e-12, atol=1e-15)
    v32 = dev.eval_device(torch.from_numpy(x.astype(np.float32)).cuda(), precision=N.PVD_FP32).cpu().numpy()
    ref32 = P.ps_water(x.astype(np.float32).astype(np.float64))
    low = ref32 < 0.5
    np.testing.assert_allclose(v32[low], ref32[low], rtol=0, atol=2e-6)


def test_ps_water_dmc_zpe_matches_the_reference_runs():
    """Plain discrete DMC on the Partridge-Schwenke surface, dt = 5 like the reference's
    tutorial runs (pyvibdmc/sample_sim_data: five runs average 4634.1 cm^-1 with a spread of
    2.2 cm^-1): 60k walkers x 4000 steps FP32 give the same ZPE within the combined error."""
    import pyvibdmc_amd as pv
    n = 60_000
    start = P.ps_eq()[None] + np.random.RandomState(5).normal(0, 0.05, (n, 3, 3))
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=4000, atoms=['H', 'H', 'O'], delta_t=5,
                     potential=pv.PartridgeSchwenkeWater(), start_structures=start, precision='fp32', seed=6,
                     write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
    sim.propagate()
    zpe = sim.zpe(onwards=1000)
    assert abs(zpe - 4634.1) < 4.0, zpe
