"""Potential managers: the plugin boundary between the DMC loop and a potential.

Two families share one interface, ``getpot(cds, timeit=False) -> (N,) hartree`` on
``(N, M, d)`` float64 walkers (reference ``simulation_utilities/potential_manager.py:71-99``):

* Host managers with the reference's constructors and behaviour, for arbitrary user
  Python potentials: ``Potential`` (process pool, :12-104), ``Potential_NoMP`` (:107-174),
  ``NN_Potential`` (:177-214), ``Potential_Direct`` (:217-253).  ``DMC_Sim`` drives them
  through the library's split path: the walkers move on the GPU, the user function runs
  on the host on the moved walkers, branching and vref run on the GPU.
* Device potentials (``DevicePotential`` subclasses) that the step kernel evaluates in
  registers, fused with the move and the branching weight.  Their ``getpot`` runs the same
  device code on a host array, so they also work anywhere the reference expects a manager.
"""
import importlib
import multiprocessing as mp
import os
import sys
import time
from itertools import repeat

import numpy as np

from . import _native as N
from .constants import Constants

__all__ = ['Potential', 'Potential_NoMP', 'NN_Potential', 'Potential_Direct', 'DevicePotential',
           'HarmonicOscillator1D', 'MorseOscillator1D', 'MorseBendTriatomic', 'TwoMonomerCluster']


def _import_function(directory, python_file, name):
    cur = os.getcwd()
    os.chdir(directory)
    try:
        sys.path.insert(0, os.getcwd())
        module = importlib.import_module(python_file.split('.')[0])
        return getattr(module, name), cur
    finally:
        os.chdir(cur)


class _Timed:
    def _finish(self, v, start, timeit):
        if getattr(self, 'pass_timestep', False):
            self.pot_kwargs['timestep'] += 1
        if timeit:
            return v, time.time() - start
        return v


class Potential(_Timed):
    """User potential in ``potential_directory/python_file``, fanned out over ``num_cores``
    pool workers with ``array_split`` / ``map`` / ``concatenate`` (potential_manager.py:12-104)."""

    def __init__(self, potential_function, potential_directory, python_file, num_cores=1,
                 pass_timestep=False, pot_kwargs=None):
        self.potential_function = potential_function
        self.python_file = python_file
        self.potential_directory = potential_directory
        self.num_cores = num_cores
        self.pass_timestep = pass_timestep
        self.pot_kwargs = pot_kwargs
        if self.pass_timestep:
            self.ct = 0
            if self.pot_kwargs is None:
                self.pot_kwargs = {'timestep': 0}
            else:
                self.pot_kwargs['timestep'] = 0
        self._init_pool()

    def _init_pot(self):
        # leaves the parent in the potential directory while the pool forks, like the
        # reference (potential_manager.py:46-57), so workers inherit it
        self._curdir = os.getcwd()
        os.chdir(self.potential_directory)
        sys.path.insert(0, os.getcwd())
        module = importlib.import_module(self.python_file.split('.')[0])
        self._pot = getattr(module, self.potential_function)

    def _init_pool(self):
        if self.num_cores <= 0:
            print('Weird number of cores specified. Defaulting to 1...')
            self.num_cores = 1
        self._init_pot()
        try:
            self._potPool = mp.get_context('fork').Pool(self.num_cores)
        finally:
            os.chdir(self._curdir)

    @property
    def pool(self):
        return self._potPool

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        if self._potPool is not None:
            chunks = np.array_split(cds, self.num_cores)
            if self.pot_kwargs is not None:
                res = self._potPool.starmap(self._pot, zip(chunks, repeat(self.pot_kwargs, len(chunks))))
            else:
                res = self._potPool.map(self._pot, chunks)
            v = np.concatenate(res)
        else:
            v = self._pot(cds)
        return self._finish(v, start, timeit)

    def mp_close(self):
        if self._potPool is not None:
            self._potPool.close()
            self._potPool.join()
            self._potPool = None

    def __getstate__(self):
        d = self.__dict__.copy()
        d['_potPool'] = None
        return d


class Potential_NoMP(_Timed):
    """User potential called in-process; ``ch_dir`` runs it from its directory
    (potential_manager.py:107-174)."""

    def __init__(self, potential_function, potential_directory, python_file, pass_timestep=False,
                 ch_dir=False, pot_kwargs=None):
        self.potential_function = potential_function
        self.python_file = python_file
        self.pass_timestep = pass_timestep
        self.potential_directory = potential_directory
        self.pot_kwargs = pot_kwargs
        if self.pass_timestep:
            if self.pot_kwargs is None:
                self.pot_kwargs = {'timestep': 0}
            else:
                self.pot_kwargs['timestep'] = 0
        self.ch_dir = ch_dir
        self._pot, self._curdir = _import_function(potential_directory, python_file, potential_function)

    def _call(self, cds):
        if self.pot_kwargs is not None:
            return self._pot(cds, self.pot_kwargs)
        return self._pot(cds)

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        if not self.ch_dir:
            v = self._call(cds)
        else:
            os.chdir(self.potential_directory)
            try:
                v = self._call(cds)
            finally:
                os.chdir(self._curdir)
        return self._finish(v, start, timeit)


class NN_Potential(Potential_NoMP):
    """``Potential_NoMP`` that also hands a model object to the function (potential_manager.py:177-214)."""

    def __init__(self, potential_function, potential_directory, python_file, model, ch_dir=False,
                 pot_kwargs=None, pass_timestep=False):
        super().__init__(potential_function, potential_directory, python_file, pass_timestep, ch_dir, pot_kwargs)
        self.model = model
        self.ch_dir = ch_dir
        self.pot_kwargs = pot_kwargs

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        if self.pot_kwargs is not None:
            v = self._pot(cds, self.model, self.pot_kwargs)
        else:
            v = self._pot(cds, self.model)
        return self._finish(v, start, timeit)


class Potential_Direct(_Timed):
    """A Python callable passed directly (potential_manager.py:217-253)."""

    def __init__(self, potential_function, pot_kwargs=None, pass_timestep=False):
        self.potential_function = potential_function
        self.pot_kwargs = pot_kwargs
        self.pass_timestep = pass_timestep
        if self.pass_timestep:
            self.ct = 0
            if self.pot_kwargs is None:
                self.pot_kwargs = {'timestep': 0}
            else:
                self.pot_kwargs['timestep'] = 0

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        if self.pot_kwargs is not None:
            v = self.potential_function(cds, self.pot_kwargs)
        else:
            v = self.potential_function(cds)
        return self._finish(v, start, timeit)


# ------------------------------------------------------------------------ device potentials
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('pyvibdmc_amd device potentials need a ROCm GPU')
    return torch


class DevicePotential:
    """A potential evaluated by the HIP library.

    ``n_atoms`` / ``n_dim`` fix the walker shape the kernels expect; ``desc`` is the
    ``pvd_pes_desc`` of the C ABI.  ``getpot`` evaluates in FP64 on the GPU."""
    n_atoms = 1
    n_dim = 3

    def __init__(self):
        self.desc = N.PesDesc()
        self.pass_timestep = False
        self.pot_kwargs = None

    @property
    def _device_pes(self):
        return self.desc

    def _check(self, cds):
        cds = np.ascontiguousarray(cds, dtype=np.float64)
        if cds.ndim != 3 or cds.shape[1] != self.n_atoms or cds.shape[2] != self.n_dim:
            raise ValueError(f'{type(self).__name__} expects walkers of shape (N, {self.n_atoms}, {self.n_dim}), '
                             f'got {cds.shape}')
        return cds

    def eval_device(self, x, precision=N.PVD_FP64):
        """V on a device tensor of shape (N, M, d) (float64 or float32); returns a tensor."""
        torch = _torch()
        x = x.contiguous()
        v = torch.empty(x.shape[0], dtype=x.dtype, device=x.device)
        N.check(N.load().pvd_pes_eval(C_ref(self.desc), precision, self.n_atoms, self.n_dim, N.ptr(x),
                                      x.shape[0], N.ptr(v), N.stream_handle()), 'pvd_pes_eval')
        return v

    def grad_device(self, x, grad_out=None, precision=N.PVD_FP64):
        """grad_out[i] * dV/dx on device (grad_out defaults to ones)."""
        torch = _torch()
        x = x.contiguous()
        go = torch.ones(x.shape[0], dtype=x.dtype, device=x.device) if grad_out is None else grad_out.contiguous()
        g = torch.empty_like(x)
        N.check(N.load().pvd_pes_vjp(C_ref(self.desc), precision, self.n_atoms, self.n_dim, N.ptr(x),
                                     x.shape[0], N.ptr(go), N.ptr(g), N.stream_handle()), 'pvd_pes_vjp')
        return g

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        torch = _torch()
        cds = self._check(cds)
        x = torch.from_numpy(cds).cuda()
        v = self.eval_device(x).cpu().numpy()
        if timeit:
            return v, time.time() - start
        return v

    def forces(self, cds):
        """-dV/dx (N, M, d) in FP64 on the GPU."""
        torch = _torch()
        cds = self._check(cds)
        x = torch.from_numpy(cds).cuda()
        return -self.grad_device(x).cpu().numpy()


def C_ref(s):
    import ctypes
    return ctypes.byref(s)


class HarmonicOscillator1D(DevicePotential):
    """V = 0.5 k x^2 on the first component of atom 0 (k = mass * omega^2).

    ``n_dim = 3`` for (N, 1, 3) walkers (only x is bound), or ``n_dim = 1`` for the
    reference's (N, 1, 1) one-dimensional walkers.  The kernel evaluates c * (x * x) with
    c = 0.5 * mass * omega**2 formed on the host exactly as NumPy evaluates
    ``0.5 * mass * omega ** 2 * x ** 2`` (harmonicOscillator1D.py:65-69)."""

    def __init__(self, k=None, mass=None, omega=None, n_dim=3):
        super().__init__()
        self.n_dim = n_dim
        if k is not None:
            c = 0.5 * k
        else:
            c = 0.5 * mass * omega ** 2
        self.k = 2 * c
        self.desc.kind = N.PES_HARMONIC_X
        self.desc.p[0] = c

    @classmethod
    def oh_stretch(cls, omega_cm=3700.):
        """The reference's ``oh_stretch_harm`` (harmonicOscillator1D.py:65-69) on (N,1,1) walkers."""
        return cls(mass=Constants.reduced_mass('O-H', to_AU=True),
                   omega=Constants.convert(omega_cm, 'wavenumbers', to_AU=True), n_dim=1)


class MorseOscillator1D(DevicePotential):
    """V = De (1 - exp(-alpha x))^2 on the first component (morse_osc_1d.py:4-12)."""

    def __init__(self, De, alpha, n_dim=1):
        super().__init__()
        self.n_dim = n_dim
        self.desc.kind = N.PES_MORSE_1D
        self.desc.p[0] = De
        self.desc.p[1] = -alpha

    @classmethod
    def oh_stretch(cls):
        """The reference's ``oh_stretch_morse`` parameters (morse_osc_1d.py:4-12)."""
        mass = Constants.reduced_mass('O-H', to_AU=True)
        omega = Constants.convert(3704.5, 'wavenumbers', to_AU=True)
        omega_x = Constants.convert(75.3, 'wavenumbers', to_AU=True)
        De = (omega ** 2 / (4 * omega_x))
        alpha = np.sqrt(mass * (omega ** 2.) / 2. / De)
        return cls(De, alpha)


class MorseBendTriatomic(DevicePotential):
    """Two Morse O-H stretches + a harmonic H-O-H bend:
    V = De (1-e^{-a(r1-re)})^2 + De (1-e^{-a(r2-re)})^2 + 0.5 kb (theta - theta0)^2.
    ``atoms`` gives the walker indices of (H1, H2, O); the default [H, H, O] order is the
    reference's water order."""
    n_atoms = 3

    def __init__(self, De=0.1916, a=1.2, re=1.81, kb=0.16, theta0=np.deg2rad(104.5), atoms=(0, 1, 2)):
        super().__init__()
        self.desc.kind = N.PES_MORSE_BEND
        for i, v in enumerate(atoms):
            self.desc.idx[i] = int(v)
        for i, v in enumerate((De, a, re, kb, theta0)):
            self.desc.p[i] = float(v)
        self.params = dict(De=De, a=a, re=re, kb=kb, theta0=theta0)


class TwoMonomerCluster(DevicePotential):
    """Two Morse+bend monomers [O,H,H,O,H,H] + O-O Morse + Coulomb between point charges
    on atoms of different monomers."""
    n_atoms = 6

    def __init__(self, De=0.1916, a=1.2, re=1.81, kb=0.16, theta0=np.deg2rad(104.5),
                 De_oo=0.008, a_oo=0.8, re_oo=5.5, q_o=-0.8, q_h=0.4):
        super().__init__()
        self.desc.kind = N.PES_CLUSTER6
        for i, v in enumerate((De, a, re, kb, theta0, De_oo, a_oo, re_oo, q_o, q_h)):
            self.desc.p[i] = float(v)


class PartridgeSchwenkeWater(DevicePotential):
    """The Partridge-Schwenke water surface (J. Chem. Phys. 106, 4618 (1997)), the reference's
    sample Fortran potential (sample_potentials/FortPots/Partridge_Schwenke_H2O/h2o_potential.py
    ``water_pot``) as a device potential.  ``atoms`` gives the walker indices of (H1, H2, O);
    the default [H, H, O] order is the reference's.  Energies only (no analytic gradient)."""
    n_atoms = 3

    def __init__(self, atoms=(0, 1, 2)):
        super().__init__()
        self.desc.kind = N.PES_PS_WATER
        for i, v in enumerate(atoms):
            self.desc.idx[i] = int(v)

    @staticmethod
    def equilibrium():
        """[H, H, O] at r_e = 0.958649 angstrom, theta_e = 104.3475 degrees (bohr)."""
        re_ = 0.958649 / 0.529177249
        t = np.deg2rad(104.3475)
        return np.array([[re_, 0.0, 0.0], [re_ * np.cos(t), re_ * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
