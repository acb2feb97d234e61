"""The fused importance-sampled step (move, Metropolis, local energy and branching in one
kernel when every tile is resident) against the oracle and against the two-kernel path.

Two environment variables read when a simulation is created steer the step for these tests:
PVD_IS_FUSE=0 turns the fused form off, and PVD_IS_SLOTS=k tiles the branching kernel (fused or
not) over at most k workgroups.  With k = 64 the fused step's population threshold is 64 tiles
x the tile size; a nominal population equal to it makes the population cross the threshold
back and forth, so a run mixes fused steps with steps whose moves k_is_move makes.  Runs with
and without fusion compared bit for bit use the same k, hence the same tile partition of
every FP64 sum.
"""
import numpy as np
import pytest

from oracle import pes as P

pytestmark = pytest.mark.gpu


def water_is_sim(pv, n, steps, precision, rng, seed_start=3):
    from tests.gpu_cases import device_trial
    start = P.water_eq()[None] + np.random.RandomState(seed_start).normal(0, 0.05, (n, 3, 3))
    return pv.DMC_Sim(num_walkers=n, num_timesteps=steps, atoms=['H', 'H', 'O'], masses=[P.M_H, P.M_H, P.M_O],
                      delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start,
                      imp_samp=device_trial({'trial': 'water_gauss'}), precision=precision, rng=rng, seed=17,
                      write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)


def run(pv, monkeypatch, fuse, slots, n, steps, precision, rng):
    monkeypatch.setenv('PVD_IS_FUSE', '1' if fuse else '0')
    monkeypatch.setenv('PVD_IS_SLOTS', str(slots or 0))
    sim = water_is_sim(pv, n, steps, precision, rng)
    monkeypatch.delenv('PVD_IS_FUSE')
    monkeypatch.delenv('PVD_IS_SLOTS')
    sim.propagate()
    coords, _ = sim._get_walkers()
    return sim, coords


def test_fused_and_two_kernel_steps_match_oracle(monkeypatch):
    """FP64, the reference's draws injected, 32 768 walkers = 64 tiles of 512: 40 steps that
    alternate between the fused and the two-kernel path match the oracle (populations,
    births and deaths exactly, vref to 1e-10, coordinates to 1e-9 as in the golden cases), and
    the run with fusion off gives bit-identical results."""
    import pyvibdmc_amd as pv
    from oracle.dmc import LegacyDraws, OracleDMC
    from tests.golden import golden_trials as gt
    from tests.golden_cases import _FnTrial
    n, T = 32_768, 40
    sim, coords = run(pv, monkeypatch, True, 64, n, T, 'fp64', LegacyDraws(31))
    pops = sim._pop_vs_tau.copy()
    assert (pops[:-1] > n).any() and (pops[:-1] <= n).any(), pops   # both paths were taken
    start = P.water_eq()[None] + np.random.RandomState(3).normal(0, 0.05, (n, 3, 3))
    orc = OracleDMC(potential=P.morse_bend, masses=[P.M_H, P.M_H, P.M_O], start_structures=start, num_walkers=n,
                    num_timesteps=T, delta_t=5, equil_steps=10 ** 9, wfn_every=10 ** 9,
                    trial=_FnTrial(gt.water_trial, gt.water_derivs), draws=LegacyDraws(31))
    orc.propagate()
    np.testing.assert_array_equal(pops, orc._pop_vs_tau)
    b, d = sim.trajectory_counts(0, T)
    np.testing.assert_array_equal(b, orc.rec['births'])
    np.testing.assert_array_equal(d, orc.rec['deaths'])
    np.testing.assert_allclose(sim._vref_vs_tau, orc._vref_vs_tau, rtol=0, atol=1e-10)
    np.testing.assert_allclose(coords, orc.coords, rtol=0, atol=1e-9)
    off, coords_off = run(pv, monkeypatch, False, 64, n, T, 'fp64', LegacyDraws(31))
    np.testing.assert_array_equal(off._vref_vs_tau, sim._vref_vs_tau)
    np.testing.assert_array_equal(coords_off, coords)


@pytest.mark.parametrize('n,cap', [(65_536, 64), (100_000, 512)])
def test_fused_fp32_production_path_is_bit_identical_to_two_kernels(monkeypatch, n, cap):
    """FP32, Philox draws, 300 steps (64-step HIP graphs): the fused step (512 tiles, every
    step fused; or 64 tiles, so that the run crosses the threshold) and the two-kernel path on
    the same tiling give bit-identical vref, populations, effective time steps and final
    walkers."""
    import pyvibdmc_amd as pv
    T = 300
    on, c_on = run(pv, monkeypatch, True, cap, n, T, 'fp32', 'philox')
    off, c_off = run(pv, monkeypatch, False, cap, n, T, 'fp32', 'philox')
    if cap:
        assert (on._pop_vs_tau[:-1] > n).any() and (on._pop_vs_tau[:-1] <= n).any()
    np.testing.assert_array_equal(on._pop_vs_tau, off._pop_vs_tau)
    np.testing.assert_array_equal(on._vref_vs_tau, off._vref_vs_tau)
    np.testing.assert_array_equal(on.eff_ts, off.eff_ts)
    np.testing.assert_array_equal(c_on, c_off)
