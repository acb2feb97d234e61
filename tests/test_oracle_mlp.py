"""Pins of the surrogate oracle: backward pass by central differences, the descriptor against
vectors the reference's DistIt wrote, its permutation invariance on the reference test's
fixtures and closed-form distances, bfloat16 rounding, Adam's first step."""
import json
import os

import numpy as np
import pytest

from oracle import mlp as O
from oracle import pes as P


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.1415926, 65504.0], np.float32)
    r = O.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert abs(r[3] + 3.140625) < 1e-7
    # every result has a zero low half-word
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'distit.npz')


def golden_cases():
    g = np.load(GOLDEN)
    for name in sorted({k.split('/')[0] for k in g.files}):
        spec = json.loads(str(g[name + '/spec']))
        yield name, g[name + '/cds'], g[name + '/zs'], spec['sorted_atoms'], spec['sorted_groups'], g[name + '/coulomb']


def coulomb_scale(zs):
    m = len(zs)
    return np.array([zs[i] * zs[j] for i in range(m) for j in range(i + 1, m)], np.float64)


def test_descriptor_matches_reference_distit_vectors():
    """Vectors the reference's DistIt(method='coulomb') wrote (tests/golden/gen_distit_golden.py):
    water and the six-atom cluster under like-atom, per-monomer, group-only and reversed-sublist
    sorting, and the reference test's trimer fixtures.  The oracle's features are those values
    divided by z_i z_j."""
    n = 0
    for name, cds, zs, sa, sg, ref in golden_cases():
        got = O.descriptor(cds, zs, sa, sg) * coulomb_scale(zs)
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0, err_msg=name)
        n += 1
    assert n == 6


def test_descriptor_permutation_invariance_of_reference_fixtures():
    """pyvibdmc/tests/test_cupy_dist.py:80-142: swapping two waters and the hydrogens of one
    leaves the monomer-sorted descriptor exactly unchanged (count_nonzero of the difference)."""
    for name, cds, zs, sa, sg, _ in golden_cases():
        if name != 'trimer_reference_fixture':
            continue
        d = O.descriptor(cds, zs, sa, sg)
        assert np.count_nonzero(np.diff(d[:4], axis=0)) == 0
        assert np.count_nonzero(np.diff(d[4:], axis=0)) == 0


def test_descriptor_closed_form_for_water():
    """[H, H, O] with the hydrogens sorted: the H nearer to O has the larger Coulomb row norm,
    so the features are (1/r_HH, max 1/r_OH, min 1/r_OH)."""
    x = P.water_eq()[None] + np.random.RandomState(0).normal(0, 0.1, (50, 3, 3))
    d = O.descriptor(x, [1, 1, 8], O.like_atoms(['H', 'H', 'O']))
    r1 = np.linalg.norm(x[:, 0] - x[:, 2], axis=1)
    r2 = np.linalg.norm(x[:, 1] - x[:, 2], axis=1)
    np.testing.assert_allclose(d[:, 0], 1 / np.linalg.norm(x[:, 0] - x[:, 1], axis=1), rtol=1e-14)
    np.testing.assert_allclose(d[:, 1], np.maximum(1 / r1, 1 / r2), rtol=1e-14)
    np.testing.assert_allclose(d[:, 2], np.minimum(1 / r1, 1 / r2), rtol=1e-14)
    np.testing.assert_array_equal(O.descriptor(x[:, [1, 0, 2]], [1, 1, 8], [[0, 1], [2]]), d)


@pytest.mark.parametrize('mistake', ['ascending', 'squared_group_norms', 'unstable_groups', 'no_group_sort'])
def test_planted_descriptor_mistakes_fail_the_golden_vectors(monkeypatch, mistake):
    real = O.atom_order

    def wrong(norms, sorted_atoms=None, sorted_groups=None):
        if mistake == 'ascending':
            return real(-norms, sorted_atoms, sorted_groups)
        if mistake == 'squared_group_norms':
            return real(norms ** 2, sorted_atoms, sorted_groups)
        if mistake == 'unstable_groups':
            return real(norms, sorted_atoms, None if sorted_groups is None else sorted_groups[::-1])
        return real(norms, sorted_atoms, None)
    monkeypatch.setattr(O, 'atom_order', wrong)
    bad = 0
    for name, cds, zs, sa, sg, ref in golden_cases():
        got = O.descriptor(cds, zs, sa, sg) * coulomb_scale(zs)
        bad += not np.allclose(got, ref, rtol=1e-12, atol=0)
    assert bad > 0


def test_backward_matches_central_differences():
    rs = np.random.RandomState(3)
    net = O.MLP(D=3, H=32, seed=4)
    xn = rs.normal(0, 1, (40, 3)).astype(np.float32)
    t = rs.normal(0, 1, 40)
    loss, _, g = net.loss_grads(xn, t, rounding=False)
    h = 1e-3
    for k in ('W0', 'b0', 'W1', 'b2', 'w3', 'b3'):
        p = net.p[k]
        flat = p.reshape(-1)
        for idx in rs.choice(flat.size, size=min(5, flat.size), replace=False):
            old = float(flat[idx])
            flat[idx] = old + h
            lp = net.loss_grads(xn, t, rounding=False)[0]
            flat[idx] = old - h
            lm = net.loss_grads(xn, t, rounding=False)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(g[k].reshape(-1)[idx] - fd) <= 2e-3 * max(1.0, abs(fd)), (k, idx, g[k].reshape(-1)[idx], fd)


def test_adam_first_step_is_signed_lr():
    p = {'a': np.array([1.0, -2.0, 3.0], np.float32)}
    opt = O.Adam(p, lr=0.01)
    opt.step(p, {'a': np.array([0.5, -4.0, 0.0], np.float32)})
    np.testing.assert_allclose(p['a'], [0.99, -1.99, 3.0], atol=1e-6)


def test_quadratic_baseline_terms_and_fit():
    """Term order 1, x_i, x_i x_j (i <= j, row order), and the least-squares fit recovers an
    exactly quadratic function of the features."""
    x = np.array([[2.0, 3.0, 5.0]])
    np.testing.assert_array_equal(O.quad_terms(x)[0], [1, 2, 3, 5, 4, 6, 10, 9, 15, 25])
    rs = np.random.RandomState(1)
    xn = rs.normal(size=(400, 3))
    c = rs.normal(size=10)
    v = O.quad_terms(xn) @ c
    np.testing.assert_allclose(O.fit_quad(xn, v), c, rtol=1e-10, atol=1e-12)
    # 15 features (six atoms): 136 terms
    assert O.quad_terms(rs.normal(size=(5, 15))).shape == (5, 136)
