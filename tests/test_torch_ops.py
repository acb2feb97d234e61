"""The torch.library operators: registration and the description round trip (CPU), and
against the direct device calls and FP64 finite differences (GPU)."""
import numpy as np
import pytest

from oracle import pes as P


def test_operators_register_and_descriptions_round_trip():
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import _native as N
    from pyvibdmc_amd import torch_ops as T
    T._register()
    for name in ('pes_energy', 'pes_vjp', 'trial_lnpsi', 'trial_vjp'):
        assert hasattr(torch.ops.pyvibdmc_amd, name)
    pot = pv.TwoMonomerCluster()
    d = T.desc_tensor(pot)
    back, m, dim = T._unpack(d, N.PesDesc)
    assert (m, dim) == (6, 3) and bytes(back) == bytes(pot.desc)
    tr = pv.water_trial()
    back, m, dim = T._unpack(T.desc_tensor(tr), N.TrialDesc)
    assert (m, dim) == (3, 3) and bytes(back) == bytes(tr.desc)
    with pytest.raises(ValueError):
        T._unpack(d[:-1], N.PesDesc)


@pytest.mark.gpu
@pytest.mark.parametrize('which', ['morse_bend', 'cluster6'])
def test_operators_match_device_calls_and_finite_differences(which):
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import torch_ops as T
    if which == 'morse_bend':
        pot, tr = pv.MorseBendTriatomic(), pv.water_trial()
        x0 = P.water_eq()[None] + np.random.RandomState(1).normal(0, 0.05, (64, 3, 3))
    else:
        pot, tr = pv.TwoMonomerCluster(), pv.cluster6_trial()
        x0 = P.cluster6_eq()[None] + np.random.RandomState(2).normal(0, 0.05, (64, 6, 3))
    x = torch.from_numpy(x0).cuda().requires_grad_(True)
    v = T.pes_energy(x, pot)
    torch.testing.assert_close(v, pot.eval_device(x.detach()), rtol=0, atol=0)
    v.sum().backward()
    torch.testing.assert_close(x.grad, pot.grad_device(x.detach()), rtol=0, atol=0)
    ln = T.trial_lnpsi(x.detach(), tr)
    torch.testing.assert_close(ln, tr.eval_device(x.detach(), want=('lnpsi',))['lnpsi'], rtol=0, atol=0)
    xs = x.detach()[:4].clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda y: T.pes_energy(y, pot), (xs,), eps=1e-6, atol=1e-6, rtol=1e-5)
    assert torch.autograd.gradcheck(lambda y: T.trial_lnpsi(y, tr), (xs,), eps=1e-6, atol=1e-6, rtol=1e-5)
