"""Importance sampling: trial wave functions and their managers.

The reference reaches psi_T through ``ImpSamp(manager)`` -> ``call_trial(cds)`` /
``call_derivs(cds)`` (simulation_utilities/imp_samp.py:16-26, imp_samp_manager.py:92-139,
197-224).  Device trial wave functions here expose the same two calls (FP64 on the GPU)
and a ``desc`` that the step kernels evaluate in registers: psi, the drift
d ln psi/dx and d2 psi/dx2 / psi at the current and the proposed walker, the Metropolis
test and the local kinetic energy, all fused into the step.

``ImpSampManager`` / ``ImpSampManager_NoMP`` keep the reference constructors for user
Python trial functions; those run on the host and are usable for checks (``call_trial``,
``call_derivs``), but the GPU step only fuses device trial functions.
"""
import ctypes
import importlib
import os
import sys

import numpy as np

from . import _native as N
from .constants import Constants

__all__ = ['ImpSamp', 'ImpSampManager', 'ImpSampManager_NoMP', 'DeviceTrial', 'HarmonicTrial1D',
           'GaussianProductTrial', 'water_trial_widths', 'TrialLogPsi', 'PotentialEnergy']


class ImpSamp:
    """Host helpers of the reference's ImpSamp (imp_samp.py:5-76)."""

    def __init__(self, imp_samp_manager):
        self.imp_manager = imp_samp_manager

    def trial(self, cds):
        return self.imp_manager.call_trial(cds)

    def drift(self, cds):
        psi_t = self.trial(cds)
        deriv, sderiv = self.imp_manager.call_derivs(cds)
        return deriv, psi_t, sderiv

    @staticmethod
    def local_kin(inv_masses_trip, sec_deriv):
        return -0.5 * np.sum(np.sum(inv_masses_trip * sec_deriv, axis=1), axis=1)

    @staticmethod
    def finite_diff(cds, trial_func):
        """3-point central differences with dx = 1e-3 bohr of psi (imp_samp.py:55-76)."""
        dx = 0.001
        num_atoms = cds.shape[1]
        num_dimz = cds.shape[-1]
        first = np.zeros(cds.shape)
        sec = np.zeros(cds.shape)
        tmp = np.zeros((len(cds), 3))
        tmp[:, 1] = trial_func(cds)
        for a in range(num_atoms):
            for d in range(num_dimz):
                cds[:, a, d] -= dx
                tmp[:, 0] = trial_func(cds)
                cds[:, a, d] += 2. * dx
                tmp[:, 2] = trial_func(cds)
                cds[:, a, d] -= dx
                first[:, a, d] = (tmp[:, 2] - tmp[:, 0]) / (2 * dx)
                sec[:, a, d] = ((tmp[:, 0] - 2. * tmp[:, 1] + tmp[:, 2]) / dx ** 2)
        return first, sec, tmp[:, 1]


def _import(directory, python_file, name):
    cur = os.getcwd()
    os.chdir(directory)
    try:
        sys.path.insert(0, os.getcwd())
        module = importlib.import_module(python_file.split('.')[0])
        return module, getattr(module, name), cur
    finally:
        os.chdir(cur)


class ImpSampManager_NoMP:
    """User trial function (and optional derivative function) called in-process
    (imp_samp_manager.py:142-224)."""

    def __init__(self, trial_function, trial_directory, python_file, chdir=False, pass_timestep=False,
                 deriv_function=None, trial_kwargs=None, deriv_kwargs=None):
        self.trial_fuc = trial_function
        self.trial_dir = trial_directory
        self.python_file = python_file
        self.pass_timestep = pass_timestep
        self.deriv_func = deriv_function
        self.trial_kwargs = trial_kwargs
        self.deriv_kwargs = deriv_kwargs
        self.chdir = chdir
        if self.pass_timestep:
            self.ct = 0
            self.trial_kwargs['timestep'] = 0
            self.deriv_kwargs['timestep'] = 0
        module, self.trial, self._curdir = _import(trial_directory, python_file, trial_function)
        if deriv_function is None:
            self.all_finite = True
            self.derivs = ImpSamp.finite_diff
        else:
            self.all_finite = False
            self.derivs = getattr(module, deriv_function)

    def call_imp_func(self, func, cds, func_kwargs=None):
        if self.chdir:
            os.chdir(self.trial_dir)
        try:
            return func(cds) if func_kwargs is None else func(cds, func_kwargs)
        finally:
            if self.chdir:
                os.chdir(self._curdir)

    def call_trial(self, cds):
        return self.call_imp_func(self.trial, cds, self.trial_kwargs)

    def call_derivs(self, cds):
        if self.all_finite:
            d, s, psi = self.derivs(cds, trial_func=self.call_trial)
            d = d / psi[:, np.newaxis, np.newaxis]
            s = s / psi[:, np.newaxis, np.newaxis]
        else:
            d, s = self.call_imp_func(self.derivs, cds, self.deriv_kwargs)
        if self.pass_timestep:
            self.ct += 1
            self.trial_kwargs['timestep'] = self.ct
            self.deriv_kwargs['timestep'] = self.ct
        return d, s


class ImpSampManager(ImpSampManager_NoMP):
    """Reference constructor of the pool-backed manager (imp_samp_manager.py:11-139).  The
    pool only parallelises host calls; here the calls run in-process."""

    def __init__(self, trial_function, trial_directory, python_file, pot_manager, pass_timestep=False,
                 new_pool_num_cores=None, deriv_function=None, trial_kwargs=None, deriv_kwargs=None):
        super().__init__(trial_function, trial_directory, python_file, chdir=False, pass_timestep=pass_timestep,
                         deriv_function=deriv_function, trial_kwargs=trial_kwargs, deriv_kwargs=deriv_kwargs)
        self.pot_manager = pot_manager
        self.nomp_pool_cores = new_pool_num_cores


# ----------------------------------------------------------------------- device trials
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('pyvibdmc_amd device trial functions need a ROCm GPU')
    return torch


class DeviceTrial:
    """psi_T evaluated by the HIP library; ``call_trial`` / ``call_derivs`` keep the manager
    contract (derivatives already divided by psi)."""
    n_atoms = 1
    n_dim = 1

    def __init__(self):
        self.desc = N.TrialDesc()

    @property
    def _device_trial(self):
        return self.desc

    def eval_device(self, x, precision=N.PVD_FP64, want=('psi', 'lnpsi', 'd1', 'd2')):
        torch = _torch()
        x = x.contiguous()
        n = x.shape[0]
        out = {}
        for k in ('psi', 'lnpsi'):
            out[k] = torch.empty(n, dtype=x.dtype, device=x.device) if k in want else None
        for k in ('d1', 'd2'):
            out[k] = torch.empty_like(x) if k in want else None
        N.check(N.load().pvd_trial_eval(ctypes.byref(self.desc), precision, self.n_atoms, self.n_dim, N.ptr(x), n,
                                        N.ptr(out['psi']), N.ptr(out['lnpsi']), N.ptr(out['d1']),
                                        N.ptr(out['d2']), N.stream_handle()), 'pvd_trial_eval')
        return out

    def vjp_device(self, x, grad_out, precision=N.PVD_FP64):
        torch = _torch()
        x = x.contiguous()
        g = torch.empty_like(x)
        N.check(N.load().pvd_trial_vjp(ctypes.byref(self.desc), precision, self.n_atoms, self.n_dim, N.ptr(x),
                                       x.shape[0], N.ptr(grad_out.contiguous()), N.ptr(g), N.stream_handle()),
                'pvd_trial_vjp')
        return g

    def _host(self, cds):
        torch = _torch()
        cds = np.ascontiguousarray(cds, dtype=np.float64)
        if cds.ndim != 3 or cds.shape[1] != self.n_atoms or cds.shape[2] != self.n_dim:
            raise ValueError(f'{type(self).__name__} expects walkers (N, {self.n_atoms}, {self.n_dim})')
        return torch.from_numpy(cds).cuda()

    def call_trial(self, cds):
        return self.eval_device(self._host(cds), want=('psi',))['psi'].cpu().numpy()

    def call_derivs(self, cds):
        o = self.eval_device(self._host(cds), want=('d1', 'd2'))
        return o['d1'].cpu().numpy(), o['d2'].cpu().numpy()


class HarmonicTrial1D(DeviceTrial):
    """psi = (alpha/pi)^(1/4) exp(-alpha x^2/2), alpha = mass * omega, on (N,1,1) walkers;
    the default parameters are those of ``harm_trial_wfn.py:12-40`` (O-H reduced mass,
    3700 cm^-1)."""

    def __init__(self, mass=None, omega=None):
        super().__init__()
        if mass is None:
            mass = Constants.reduced_mass('O-H')
        if omega is None:
            omega = Constants.convert(3700, 'wavenumbers', to_AU=True)
        alpha = mass * omega
        self.alpha = alpha
        self.desc.kind = N.TRIAL_HARMONIC_1D
        self.desc.c0 = (alpha / np.pi) ** 0.25
        self.desc.c1 = alpha
        self.desc.c2 = alpha ** 2


class GaussianProductTrial(DeviceTrial):
    """psi_T = exp(c0) * prod_k exp(-alpha_k (q_k - q0_k)^2 / 2) over bond lengths
    ``bonds = [(i, j, alpha, r0)]`` and bond angles ``angles = [(i, j, k, alpha, theta0)]``
    (vertex j), with analytic Cartesian derivatives including every cross term of the
    product (the chain rule of su/imp_samp_helper.py, for any number of coordinates)."""
    n_dim = 3

    def __init__(self, n_atoms, bonds, angles=(), log_norm=0.0):
        super().__init__()
        if len(bonds) > 8 or len(angles) > 4:
            raise ValueError('at most 8 bonds and 4 angles')
        self.n_atoms = n_atoms
        d = self.desc
        d.kind = N.TRIAL_GAUSS
        d.nb = len(bonds)
        d.na = len(angles)
        for k, (i, j, a, r0) in enumerate(bonds):
            d.bi[k], d.bj[k], d.balpha[k], d.br0[k] = int(i), int(j), float(a), float(r0)
        for k, (i, j, kk, a, t0) in enumerate(angles):
            d.ai[k], d.aj[k], d.ak[k], d.aalpha[k], d.ath0[k] = int(i), int(j), int(kk), float(a), float(t0)
        d.c0 = float(log_norm)
        self.bonds, self.angles = list(bonds), list(angles)


def water_trial_widths(m_h=1837.15, m_o=29156.95, De=0.1916, a=1.2, re=1.81, kb=0.16, theta0=np.deg2rad(104.5)):
    """Gaussian widths from the harmonic frequencies of the Morse+bend surface:
    alpha_r = mu_OH omega_r with omega_r = a sqrt(2 De / mu_OH); alpha_theta = omega_theta / G_thth
    with omega_theta = sqrt(kb G_thth) and G_thth of the bent triatomic (call_trl_h2o.py:38-46)."""
    mu = m_h * m_o / (m_h + m_o)
    w_r = a * np.sqrt(2.0 * De / mu)
    g = 2.0 / (m_h * re ** 2) + (2.0 / re ** 2 - 2.0 * np.cos(theta0) / re ** 2) / m_o
    w_t = np.sqrt(kb * g)
    return dict(alpha_r=mu * w_r, alpha_t=w_t / g, omega_r=w_r, omega_t=w_t, g_tt=g, mu=mu)


def water_trial(h1=0, h2=1, o=2, re=1.81, theta0=np.deg2rad(104.5), **kw):
    """Product of Gaussians in both O-H stretches and the bend of an [H, H, O] walker."""
    p = water_trial_widths(re=re, theta0=theta0, **kw)
    return GaussianProductTrial(3, bonds=[(h1, o, p['alpha_r'], re), (h2, o, p['alpha_r'], re)],
                                angles=[(h1, o, h2, p['alpha_t'], theta0)])


def cluster6_trial(re=1.81, theta0=np.deg2rad(104.5)):
    """Per-monomer Gaussian product for the [O,H,H,O,H,H] cluster (6 internal coordinates)."""
    p = water_trial_widths(re=re, theta0=theta0)
    ar, at = p['alpha_r'], p['alpha_t']
    return GaussianProductTrial(6, bonds=[(1, 0, ar, re), (2, 0, ar, re), (4, 3, ar, re), (5, 3, ar, re)],
                                angles=[(1, 0, 2, at, theta0), (4, 3, 5, at, theta0)])


# ------------------------------------------------------------------- autograd wrappers
def _autograd():
    import torch

    class TrialLogPsi(torch.autograd.Function):
        """ln psi_T(x) with a hand-written backward kernel: the backward recomputes the
        forward at x and returns grad_out * d ln psi / dx (the drift direction)."""

        @staticmethod
        def forward(ctx, x, trial):
            ctx.save_for_backward(x)
            ctx.trial = trial
            prec = N.PVD_FP64 if x.dtype == torch.float64 else N.PVD_FP32
            return trial.eval_device(x, precision=prec, want=('lnpsi',))['lnpsi']

        @staticmethod
        def backward(ctx, g):
            (x,) = ctx.saved_tensors
            prec = N.PVD_FP64 if x.dtype == torch.float64 else N.PVD_FP32
            return ctx.trial.vjp_device(x, g, precision=prec), None

    class PotentialEnergy(torch.autograd.Function):
        """V(x) of a device potential with a hand-written backward kernel (dV/dx; forces are
        its negative)."""

        @staticmethod
        def forward(ctx, x, pot):
            ctx.save_for_backward(x)
            ctx.pot = pot
            prec = N.PVD_FP64 if x.dtype == torch.float64 else N.PVD_FP32
            return pot.eval_device(x, precision=prec)

        @staticmethod
        def backward(ctx, g):
            (x,) = ctx.saved_tensors
            prec = N.PVD_FP64 if x.dtype == torch.float64 else N.PVD_FP32
            return ctx.pot.grad_device(x, grad_out=g, precision=prec), None

    return TrialLogPsi, PotentialEnergy


def __getattr__(name):
    if name in ('TrialLogPsi', 'PotentialEnergy'):
        t, p = _autograd()
        globals()['TrialLogPsi'] = t
        globals()['PotentialEnergy'] = p
        return globals()[name]
    raise AttributeError(name)
