"""The device potentials and trial wave functions as PyTorch operators.

``torch.library`` custom operators over the library's C entry points, with autograd through
the hand-written vector-Jacobian kernels (the potential's forces, the trial's drift):

    torch.ops.pyvibdmc_amd.pes_energy(x, desc)       -> V(x)                  (n,)
    torch.ops.pyvibdmc_amd.pes_vjp(x, desc, g)       -> g_i dV/dx_i           (n, M, d)
    torch.ops.pyvibdmc_amd.trial_lnpsi(x, desc)      -> ln psi_T(x)           (n,)
    torch.ops.pyvibdmc_amd.trial_vjp(x, desc, g)     -> g_i d ln psi_T/dx_i   (n, M, d)

``x`` is a contiguous (n, M, d) float32 or float64 tensor on the GPU; ``desc`` is the C
description of the potential (``pvd_pes_desc``) or trial (``pvd_trial_desc``) as a CPU uint8
tensor, from ``desc_tensor(potential_or_trial)``.  The reference evaluates its potentials
through ``Potential.getpot`` (simulation_utilities/potential_manager.py:71-99) and its trial
derivatives through ``ImpSampManager.call_derivs`` (simulation_utilities/imp_samp_manager.py:
92-139); these operators are the same evaluations as differentiable tensor functions.
"""
import ctypes

from . import _native as N

__all__ = ['desc_tensor', 'pes_energy', 'trial_lnpsi']

_REGISTERED = False


def desc_tensor(obj):
    """The C description of a device potential or trial as a CPU uint8 tensor, with the atom
    and dimension counts appended (two int32)."""
    import numpy as np
    import torch
    raw = bytes(obj.desc) + np.array([obj.n_atoms, obj.n_dim], np.int32).tobytes()
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8)


def _unpack(desc, cls):
    raw = desc.cpu().numpy().tobytes()
    n = ctypes.sizeof(cls)
    if len(raw) != n + 8:
        raise ValueError(f'description of {len(raw)} bytes, expected {n + 8}')
    import numpy as np
    m, d = np.frombuffer(raw[n:], np.int32)
    return cls.from_buffer_copy(raw[:n]), int(m), int(d)


def _prec(x):
    import torch
    if x.dtype == torch.float64:
        return N.PVD_FP64
    if x.dtype == torch.float32:
        return N.PVD_FP32
    raise TypeError('walkers must be float32 or float64')


def _register():
    global _REGISTERED
    if _REGISTERED:
        return
    import torch

    @torch.library.custom_op('pyvibdmc_amd::pes_energy', mutates_args=())
    def pes_energy_op(x: torch.Tensor, desc: torch.Tensor) -> torch.Tensor:
        pes, m, d = _unpack(desc, N.PesDesc)
        x = x.contiguous()
        v = torch.empty(x.shape[0], dtype=x.dtype, device=x.device)
        N.check(N.load().pvd_pes_eval(ctypes.byref(pes), _prec(x), m, d, N.ptr(x), x.shape[0], N.ptr(v),
                                      N.stream_handle()), 'pvd_pes_eval')
        return v

    @torch.library.custom_op('pyvibdmc_amd::pes_vjp', mutates_args=())
    def pes_vjp_op(x: torch.Tensor, desc: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
        pes, m, d = _unpack(desc, N.PesDesc)
        x = x.contiguous()
        out = torch.empty_like(x)
        N.check(N.load().pvd_pes_vjp(ctypes.byref(pes), _prec(x), m, d, N.ptr(x), x.shape[0],
                                     N.ptr(g.contiguous()), N.ptr(out), N.stream_handle()), 'pvd_pes_vjp')
        return out

    @torch.library.custom_op('pyvibdmc_amd::trial_lnpsi', mutates_args=())
    def trial_lnpsi_op(x: torch.Tensor, desc: torch.Tensor) -> torch.Tensor:
        tr, m, d = _unpack(desc, N.TrialDesc)
        x = x.contiguous()
        ln = torch.empty(x.shape[0], dtype=x.dtype, device=x.device)
        N.check(N.load().pvd_trial_eval(ctypes.byref(tr), _prec(x), m, d, N.ptr(x), x.shape[0], None, N.ptr(ln),
                                        None, None, N.stream_handle()), 'pvd_trial_eval')
        return ln

    @torch.library.custom_op('pyvibdmc_amd::trial_vjp', mutates_args=())
    def trial_vjp_op(x: torch.Tensor, desc: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
        tr, m, d = _unpack(desc, N.TrialDesc)
        x = x.contiguous()
        out = torch.empty_like(x)
        N.check(N.load().pvd_trial_vjp(ctypes.byref(tr), _prec(x), m, d, N.ptr(x), x.shape[0],
                                       N.ptr(g.contiguous()), N.ptr(out), N.stream_handle()), 'pvd_trial_vjp')
        return out

    for op in (pes_energy_op, trial_lnpsi_op):
        op.register_fake(lambda x, desc: x.new_empty(x.shape[0]))
    for op in (pes_vjp_op, trial_vjp_op):
        op.register_fake(lambda x, desc, g: torch.empty_like(x))

    def setup(ctx, inputs, output):
        ctx.save_for_backward(inputs[0], inputs[1])

    def pes_backward(ctx, grad):
        x, desc = ctx.saved_tensors
        return torch.ops.pyvibdmc_amd.pes_vjp(x, desc, grad), None

    def trial_backward(ctx, grad):
        x, desc = ctx.saved_tensors
        return torch.ops.pyvibdmc_amd.trial_vjp(x, desc, grad), None

    torch.library.register_autograd('pyvibdmc_amd::pes_energy', pes_backward, setup_context=setup)
    torch.library.register_autograd('pyvibdmc_amd::trial_lnpsi', trial_backward, setup_context=setup)
    _REGISTERED = True


def pes_energy(x, potential):
    """V(x) of a device potential as a differentiable tensor function (backward: dV/dx)."""
    _register()
    import torch
    return torch.ops.pyvibdmc_amd.pes_energy(x, desc_tensor(potential))


def trial_lnpsi(x, trial):
    """ln psi_T(x) of a device trial as a differentiable tensor function (backward: the drift
    direction d ln psi_T / dx)."""
    _register()
    import torch
    return torch.ops.pyvibdmc_amd.trial_lnpsi(x, desc_tensor(trial))
