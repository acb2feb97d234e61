"""Independent pins of three oracle functions whose values nothing else fixes.

* ``pes.cluster6``: a value assembled by hand at a stated geometry (both monomers at their
  equilibrium, monomer B straight above A), where every inter-monomer distance is one of
  four closed forms; and the separated-monomer limit, where the Coulomb sum between two
  neutral monomers decays as the dipole-dipole energy mu_A . mu_B / R^3 (dipoles
  perpendicular to the separation).
* ``mlp.MLP.forward`` / ``loss_grads``: against torch's float64 linear layers, SiLU and
  autograd on the CPU (without the bfloat16 rounding), and the forward with rounding
  against torch's own bfloat16 cast of the same operands.
* ``mlp.Adam``: ten steps against ``torch.optim.Adam`` on float32 parameters.

Each check fails on the mistakes named in its docstring (planted on the oracle side in
``test_pins_catch_planted_mistakes``).
"""
import numpy as np
import pytest

from oracle import mlp as O
from oracle import pes as P

torch = pytest.importorskip('torch')

# ---------------------------------------------------------------------------- cluster6
R_SEP = 6.25  # O-O separation of the stated geometry (bohr), away from re_OO = 5.5


# the configuration's numbers, typed here rather than read from the oracle module
RE, THETA0 = 1.81, np.deg2rad(104.5)
DE_OO, A_OO, RE_OO = 0.008, 0.8, 5.5
Q_O, Q_H = -0.8, 0.4


def stated_cluster(R):
    """[O,H,H,O,H,H]: monomer A at equilibrium in the xy plane (O at the origin), monomer B
    the same translated by R along z."""
    a = np.array([[0.0, 0.0, 0.0], [RE, 0.0, 0.0], [RE * np.cos(THETA0), RE * np.sin(THETA0), 0.0]])
    return np.concatenate([a, a + [0.0, 0.0, R]])[None]


def hand_value(R):
    """Both monomers at equilibrium: no intramolecular energy.  Inter-monomer distances:
    O-O and H1-H1', H2-H2' are R; O-H' and H-O' are sqrt(re^2 + R^2); H1-H2' and H2-H1'
    are sqrt(d_HH^2 + R^2) with d_HH = 2 re sin(theta0 / 2)."""
    qoo, qoh, qhh = Q_O * Q_O, Q_O * Q_H, Q_H * Q_H
    d_hh = 2.0 * RE * np.sin(THETA0 / 2.0)
    r_oh = np.hypot(RE, R)
    r_hh = np.hypot(d_hh, R)
    coulomb = qoo / R + 4.0 * qoh / r_oh + 2.0 * qhh / R + 2.0 * qhh / r_hh
    morse_oo = DE_OO * (1.0 - np.exp(-A_OO * (R - RE_OO))) ** 2
    return coulomb + morse_oo


def test_cluster6_hand_value_at_stated_geometry():
    """Fails on a flipped Coulomb sign, an intra-monomer pair in the Coulomb list, a
    wrong O-O Morse parameter or a charge swapped between O and H."""
    v = P.cluster6(stated_cluster(R_SEP))[0]
    # the same number typed out (evaluated once from hand_value, 17 digits)
    assert np.isclose(hand_value(R_SEP), 0.005062213180592488, rtol=0, atol=1e-15)
    assert abs(v - hand_value(R_SEP)) < 1e-15


def test_cluster6_separated_monomers():
    """R -> infinity: V -> V_A + V_B + De_OO, the Coulomb sum vanishing like the
    dipole-dipole energy mu_A . mu_B / R^3 of two neutral monomers."""
    rs = np.random.RandomState(5)
    # a distorted [O, H, H] monomer in the xy plane, so its dipole is perpendicular to z
    mono = P.water_eq()[[2, 0, 1]] + rs.normal(0, 0.05, (3, 3))
    mono[:, 2] = 0.0
    q = np.array([Q_O, Q_H, Q_H])
    assert abs(q.sum()) < 1e-15
    mu = (q[:, None] * mono).sum(axis=0)
    va = P.morse_bend(mono[None], h1=1, h2=2, o=0)[0]
    for R in (200.0, 400.0):
        cds = np.concatenate([mono, mono + [0.0, 0.0, R]])[None]
        inter = P.cluster6(cds)[0] - 2 * va - DE_OO * (1.0 - np.exp(-A_OO * (R - RE_OO))) ** 2
        assert np.isclose(inter * R ** 3, mu @ mu, rtol=2e-3), (R, inter * R ** 3, mu @ mu)
    far = np.concatenate([mono, mono + [0.0, 0.0, 1e6]])[None]
    assert abs(P.cluster6(far)[0] - (2 * va + DE_OO)) < 1e-12


# ----------------------------------------------------------------------------- network
def torch_net(net):
    p = {k: torch.tensor(v.astype(np.float64), requires_grad=True) for k, v in net.p.items()}
    return p


def torch_forward(p, x, cast=None):
    c = cast or (lambda a: a)
    a = c(x)
    for l in range(3):
        z = torch.nn.functional.linear(a, c(p[f'W{l}']), p[f'b{l}'])
        h = torch.nn.functional.silu(z)
        a = c(h.float()).double() if (cast is not None and l < 2) else h
    return a @ p['w3'] + p['b3'][0]


@pytest.mark.parametrize('D,H', [(3, 32), (15, 64)])
def test_mlp_forward_and_gradients_match_torch(D, H):
    """Fails on a transposed W1 (or any swapped weight), a dropped bias or a wrong
    activation derivative."""
    rs = np.random.RandomState(D)
    net = O.MLP(D=D, H=H, seed=D + 1)
    for k in ('b0', 'b1', 'b2', 'b3'):
        net.p[k] = rs.normal(0, 0.2, net.p[k].shape).astype(np.float32)
    xn = rs.normal(0, 1, (64, D)).astype(np.float32)
    t = rs.normal(0, 1, 64)
    y, _ = net.forward(xn, rounding=False)
    p = torch_net(net)
    yt = torch_forward(p, torch.tensor(xn.astype(np.float64)))
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-13, atol=1e-13)
    loss, _, g = net.loss_grads(xn, t, rounding=False)
    lt = torch.mean((yt - torch.tensor(t)) ** 2)
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12 * max(1.0, abs(loss))
    for k, v in p.items():
        np.testing.assert_allclose(g[k].reshape(v.shape), v.grad.numpy(), rtol=1e-10, atol=1e-12, err_msg=k)


def test_mlp_forward_rounding_points_match_torch_bfloat16():
    """The oracle's bfloat16 operands (inputs and hidden activations of the three MFMA
    layers, and their weights) equal torch's bfloat16 cast of the same values."""
    rs = np.random.RandomState(7)
    net = O.MLP(D=3, H=128, seed=8)
    xn = rs.normal(0, 1, (256, 3)).astype(np.float32)
    y, _ = net.forward(xn, rounding=True)
    p = torch_net(net)

    def cast(a):
        return a.float().to(torch.bfloat16).double()
    yt = torch_forward(p, torch.tensor(xn.astype(np.float64)), cast=cast)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)


# -------------------------------------------------------------------------------- Adam
def test_adam_ten_steps_match_torch():
    """Fails on a bias correction frozen at t = 1, swapped betas or eps inside the root."""
    rs = np.random.RandomState(11)
    shapes = dict(a=(7,), W=(5, 3))
    params = {k: rs.normal(0, 1, s).astype(np.float32) for k, s in shapes.items()}
    tp = [torch.tensor(params[k].copy(), requires_grad=True) for k in shapes]
    opt_t = torch.optim.Adam(tp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, foreach=False)
    opt = O.Adam(params, lr=1e-3)
    for step in range(10):
        grads = {k: (rs.normal(0, 1, s) * 10.0 ** rs.uniform(-3, 1)).astype(np.float32) for k, s in shapes.items()}
        for t_, k in zip(tp, shapes):
            t_.grad = torch.tensor(grads[k])
        opt_t.step()
        opt.step(params, grads)
        for t_, k in zip(tp, shapes):
            np.testing.assert_allclose(params[k], t_.detach().numpy(), rtol=2e-6, atol=2e-7,
                                       err_msg=f'{k} step {step + 1}')


# ------------------------------------------------------------- the checks catch mistakes
def test_pins_catch_planted_mistakes(monkeypatch):
    """Each pin above rejects a plausible mistake planted in the oracle."""
    geo = stated_cluster(R_SEP)
    good = P.cluster6(geo)[0]
    morse_oo = DE_OO * (1.0 - np.exp(-A_OO * (R_SEP - RE_OO))) ** 2
    coulomb = good - morse_oo  # both monomers at equilibrium: no intramolecular energy
    # flipped Coulomb sign: off by twice the Coulomb sum
    assert abs((morse_oo - coulomb) - hand_value(R_SEP)) > 1e-3
    # an intra-monomer O-H pair in the Coulomb list
    assert abs((good + Q_O * Q_H / RE) - hand_value(R_SEP)) > 1e-2
    # wrong O-O Morse range parameter
    monkeypatch.setattr(P, 'OO_A', 0.9)
    assert abs(P.cluster6(geo)[0] - hand_value(R_SEP)) > 1e-6
    monkeypatch.undo()
    # charges swapped between O and H
    monkeypatch.setattr(P, 'Q_O', 0.4)
    monkeypatch.setattr(P, 'Q_H', -0.8)
    assert abs(P.cluster6(geo)[0] - hand_value(R_SEP)) > 1e-3
    monkeypatch.undo()
    # transposed W1 changes the forward
    net = O.MLP(D=3, H=16, seed=2)
    xn = np.random.RandomState(0).normal(0, 1, (8, 3)).astype(np.float32)
    y, _ = net.forward(xn, rounding=False)
    net.p['W1'] = net.p['W1'].T.copy()
    y2, _ = net.forward(xn, rounding=False)
    assert np.max(np.abs(y - y2)) > 1e-6

    # Adam with the bias correction frozen at t = 1 drifts from the correct one by t = 3
    class Frozen(O.Adam):
        def step(self, params, grads):
            self.t = 0
            super().step(params, grads)
    p1 = {'a': np.ones(4, np.float32)}
    p2 = {'a': np.ones(4, np.float32)}
    a1, a2 = O.Adam(p1), Frozen(p2)
    for _ in range(3):
        g = {'a': np.array([1.0, -0.5, 0.25, 2.0], np.float32)}
        a1.step(p1, g)
        a2.step(p2, g)
    assert np.max(np.abs(p1['a'] - p2['a'])) > 1e-5
