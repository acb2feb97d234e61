"""The oracle reproduces the reference's own seeded trajectories (golden vectors).

Bit-exact: populations, per-walker branching counts / clone sources, births, deaths,
acceptance ratios, final coordinates.  vref within 1e-12 hartree (same NumPy ops; the
tolerance only absorbs summation-order noise).
"""
import numpy as np
import pytest

from tests.golden_cases import CASES, load, oracle_for, split_counts, split_src


@pytest.mark.parametrize('name', CASES)
def test_oracle_matches_reference(name):
    g = load(name)
    o = oracle_for(name, g)
    np.testing.assert_array_equal(o.coords, g['start'])
    o.propagate()
    T = int(g['num_timesteps'])
    assert np.array_equal(np.array(o.rec['pop_in']), g['pop_in'])
    if 'counts_flat' in g:
        for t, c in enumerate(split_counts(g)):
            idx = o.rec['idx'][t]
            np.testing.assert_array_equal(np.bincount(idx, minlength=len(c)), c, err_msg=f'step {t}')
        np.testing.assert_array_equal(o.rec['births'], g['births'])
        np.testing.assert_array_equal(o.rec['deaths'], g['deaths'])
    else:
        for t, s in enumerate(split_src(g)):
            np.testing.assert_array_equal(o.rec['src'][t], s, err_msg=f'step {t}')
        np.testing.assert_array_equal(o.rec['n_branched'], g['n_branched'])
        np.testing.assert_array_equal(o.cont_wts, g['final_wts'])
    if 'dt_factor' in g:
        np.testing.assert_array_equal(np.array(o.rec['accept']) / np.array(g['pop_in']), g['dt_factor'])
        np.testing.assert_allclose(o.eff_ts, g['eff_ts'], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o._vref_vs_tau, g['vref'], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o._pop_vs_tau, g['pop'], rtol=1e-13, atol=0)
    assert len(o._vref_vs_tau) == T
    np.testing.assert_array_equal(o.coords, g['final_coords'])
    np.testing.assert_allclose(o.V, g['final_pots'], rtol=0, atol=1e-12)
    # descendant weighting windows
    assert len(o.wfns) == int(g['n_wfn'])
    for i, (step, parent, dw, pw) in enumerate(o.wfns):
        assert step == int(g[f'wfn{i}_step'])
        np.testing.assert_array_equal(parent, g[f'wfn{i}_coords'])
        np.testing.assert_allclose(dw, g[f'wfn{i}_desc_wts'], rtol=1e-14, atol=0)
        if pw is not None:
            np.testing.assert_array_equal(pw, g[f'wfn{i}_parent_wts'])


def test_golden_cases_exercise_branching():
    """The fixtures are only useful if they contain births, deaths, splits and ties."""
    g = load('c0_discrete')
    assert g['births'].sum() > 1000 and g['deaths'].sum() > 1000
    c = load('c1_continuous')
    assert c['n_branched'].sum() > 200
    d = load('c0_desc')
    assert int(d['n_wfn']) >= 2
