"""Pins of the surrogate oracle: backward pass by central differences, descriptor
invariances and closed-form distances, bfloat16 rounding, Adam's first step."""
import numpy as np

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


def test_descriptor_closed_form_and_permutation_invariance():
    x = P.water_eq()[None] + np.random.RandomState(0).normal(0, 0.1, (50, 3, 3))
    pairs, sizes = O.pair_groups(['H', 'H', 'O'])
    assert pairs == [(0, 1), (0, 2), (1, 2)] and sizes == [1, 2]
    d = O.descriptor(x, pairs, sizes)
    r1 = np.linalg.norm(x[:, 0] - x[:, 2], axis=1)
    r2 = np.linalg.norm(x[:, 1] - x[:, 2], axis=1)
    np.testing.assert_allclose(d[:, 1], np.maximum(1 / r1, 1 / r2), rtol=1e-14)
    np.testing.assert_allclose(d[:, 2], np.minimum(1 / r1, 1 / r2), rtol=1e-14)
    swapped = x[:, [1, 0, 2]]
    np.testing.assert_array_equal(O.descriptor(swapped, pairs, sizes), d)
    # six atoms: swapping the monomers and the hydrogens inside one leaves it unchanged
    c = P.cluster6_eq()[None] + np.random.RandomState(1).normal(0, 0.1, (20, 6, 3))
    pairs6, sizes6 = O.pair_groups(['O', 'H', 'H', 'O', 'H', 'H'])
    assert len(pairs6) == 15 and sorted(sizes6) == [1, 6, 8]
    d6 = O.descriptor(c, pairs6, sizes6)
    np.testing.assert_allclose(O.descriptor(c[:, [3, 4, 5, 0, 2, 1]], pairs6, sizes6), d6, rtol=1e-14)


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
