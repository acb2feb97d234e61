"""The allocation-free gradient entry point (pvd_mlp_grads_noalloc) against pvd_mlp_grads
with a caller workspace: the same kernels on a static scratch array, so the gradients, the
network outputs and the loss partials are identical bit for bit.  Two calls in a row on one
device, the second with a smaller batch (the scratch keeps the first call's contents), and
a batch whose workspace exceeds the scratch is rejected."""
import ctypes

import numpy as np
import pytest

from oracle import pes as P

pytestmark = pytest.mark.gpu


def _walkers(n, seed):
    return P.water_eq()[None] + np.random.RandomState(seed).normal(0, 0.08, (n, 3, 3))


@pytest.mark.parametrize('elements,hidden,sizes', [(['H', 'H', 'O'], 128, (8192, 3000)),
                                                   (['O', 'H', 'H', 'O', 'H', 'H'], 256, (4096, 1000))])
def test_noalloc_grads_equal_workspace_grads(elements, hidden, sizes):
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import _native as N
    from pyvibdmc_amd.surrogate import _lib
    lib = _lib()
    m = pv.SurrogateModel(elements, hidden=hidden, seed=3)
    big = _walkers(sizes[0], 1) if len(elements) == 3 else P.cluster6_eq()[None] + np.random.RandomState(1).normal(0, 0.08, (sizes[0], 6, 3))
    pes = P.morse_bend if len(elements) == 3 else P.cluster6
    m.set_normalisation(big, pes(big))
    for n in sizes:  # the second call is smaller: it reuses the scratch the first one filled
        x = big[:n]
        e = pes(x)
        xt = m.features(x, normalise=True)
        t = torch.from_numpy(((e - m.desc.ymean) / m.desc.ystd).astype(np.float32)).cuda()
        assert lib.pvd_mlp_workspace_bytes(m.D, m.H, n) <= lib.pvd_mlp_noalloc_scratch_bytes()
        loss_ref, y_ref = m.loss_and_grads(xt, t)
        g_ref = m.grads.clone()
        lp_ref = m._lp[:(n + 255) // 256].clone()
        g = torch.full_like(m.grads, float('nan'))
        y = torch.full((n,), float('nan'), dtype=torch.float32, device='cuda')
        lp = torch.full(((n + 255) // 256,), float('nan'), dtype=torch.float64, device='cuda')
        rc = lib.pvd_mlp_grads_noalloc(ctypes.byref(m.desc), xt.data_ptr(), t.data_ptr(), n, g.data_ptr(),
                                       y.data_ptr(), lp.data_ptr(), N.stream_handle())
        assert rc == 0, lib.pvd_mlp_last_error()
        torch.cuda.synchronize()
        assert torch.equal(g, g_ref)
        assert torch.equal(y, y_ref)
        assert torch.equal(lp, lp_ref)


def test_noalloc_rejects_a_batch_beyond_the_scratch():
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import _native as N
    from pyvibdmc_amd.surrogate import _lib
    lib = _lib()
    m = pv.SurrogateModel(['H', 'H', 'O'], hidden=128, seed=3)
    n = 1
    while lib.pvd_mlp_workspace_bytes(m.D, m.H, n) <= lib.pvd_mlp_noalloc_scratch_bytes():
        n *= 2
    dummy = torch.zeros(4, dtype=torch.float32, device='cuda')
    rc = lib.pvd_mlp_grads_noalloc(ctypes.byref(m.desc), dummy.data_ptr(), dummy.data_ptr(), n, dummy.data_ptr(),
                                   None, None, N.stream_handle())
    assert rc == 7
    assert b'static scratch' in lib.pvd_mlp_last_error()
