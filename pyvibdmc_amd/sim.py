"""DMC_Sim and dmc_restart: the reference's engine surface over the HIP step kernels.

Mirrors ``pyvibdmc/pyvibdmc.py`` (``DMC_Sim.__init__`` :67-130, ``_initialize`` :132-297,
``propagate`` :701-876, ``run`` :878-932, ``dmc_restart`` :949-965).  The walkers live on
the GPU for the whole run; the host only wakes up at checkpoints, descendant-weighting
window edges and training dumps, and otherwise issues one fused kernel step after another
(``pvd_sim_run``).

Draw sources (``rng``):
* ``'philox'`` (default): counter-based Philox4x32-10 on the device, keyed by ``seed`` and
  addressed by (walker slot, step) - restarts are exactly reproducible.
* ``'numpy'``: the reference's own draw stream, the global legacy ``np.random`` in the
  reference's order (normals of shape (N, d, M) then uniforms, pyvibdmc.py:392, 544-546, 601),
  one step per launch.  With ``precision='fp64'`` a run seeded with ``np.random.seed`` follows
  the reference's trajectory.
* any object with ``displacement(sigmas, shape)`` and ``uniform(n)``: same, from that object.
"""
import copy
import os
import pickle
import time

import numpy as np

from . import _native as N
from . import hdf5
from .constants import Constants, get_atomic_num

__all__ = ['DMC_Sim', 'dmc_restart']


class _NumpyDraws:
    """The reference's global-np.random draw order."""

    @staticmethod
    def displacement(sigmas, shape):
        n, m, d = shape
        return np.random.normal(0.0, sigmas, size=(n, d, m)).transpose(0, 2, 1)

    @staticmethod
    def uniform(n):
        return np.random.random(n)


class _Logger:
    def __init__(self, fname, overwrite=False, enabled=True):
        self.fl = open(fname, 'w' if overwrite else 'a') if enabled else None

    def write(self, s):
        if self.fl:
            self.fl.write(s)

    def close(self):
        if self.fl:
            self.fl.close()
            self.fl = None


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('pyvibdmc_amd.DMC_Sim runs on a ROCm GPU; none is visible')
    return torch


class DMC_Sim:
    """Diffusion Monte Carlo simulation on one MI355X.

    Keyword arguments are the reference's (pyvibdmc.py:67-98) plus:
    ``precision`` ('fp64' | 'fp32': FP32 coordinates and per-walker PES, FP64 weights and
    reductions), ``rng`` (see the module docstring), ``seed`` (Philox key),
    ``write_files`` (False keeps everything in memory), ``capacity`` (walker slots)."""

    def __init__(self, sim_name="DMC_Sim", output_folder="exSimResults", weighting='discrete', num_walkers=10000,
                 num_timesteps=20000, equil_steps=2000, chkpt_every=1000, wfn_every=1000, desc_wt_steps=100,
                 atoms=[], delta_t=1, potential=None, masses=None, start_structures=None, start_cont_wts=None,
                 branch_every=1, log_every=1, cur_timestep=0, cont_wt_thresh=None, imp_samp=None,
                 imp_samp_oned=False, second_impsamp_displacement=False, excited_state_imp_samp=False,
                 adiabatic_dmc=None, fixed_node=None, DEBUG_alpha=None, DEBUG_save_desc_wt_tracker=None,
                 DEBUG_save_training_every=None, DEBUG_save_before_bod=False, DEBUG_mass_change=None,
                 precision='fp64', rng='philox', seed=None, write_files=True, capacity=0):
        unsupported = dict(second_impsamp_displacement=second_impsamp_displacement,
                           excited_state_imp_samp=excited_state_imp_samp, adiabatic_dmc=adiabatic_dmc is not None,
                           fixed_node=fixed_node is not None, DEBUG_mass_change=DEBUG_mass_change is not None,
                           DEBUG_save_desc_wt_tracker=bool(DEBUG_save_desc_wt_tracker))
        bad = [k for k, v in unsupported.items() if v]
        if bad:
            raise NotImplementedError(f'not part of the GPU propagation path: {", ".join(bad)}')
        self.atoms = atoms
        self.sim_name = sim_name
        self.output_folder = output_folder
        self.num_walkers = num_walkers
        self.num_timesteps = num_timesteps
        self.potential_info = dict(vars(potential)) if potential is not None else {}
        self.potential_obj = potential
        self.potential = potential.getpot if potential is not None else None
        self.weighting = weighting.lower()
        self.desc_wt_time_steps = desc_wt_steps
        self.branch_every = branch_every
        self.delta_t = delta_t
        self.start_structures = start_structures
        self.start_cont_wts = start_cont_wts
        self.masses = masses
        self.equil_steps = equil_steps
        self.chkpt_every = chkpt_every
        self.wfn_every = wfn_every
        self.log_every = log_every
        self.cur_timestep = cur_timestep
        self.cont_wt_thresh = cont_wt_thresh
        self.impsamp_manager = imp_samp
        self.imp1d = imp_samp_oned
        self._deb_training_every = DEBUG_save_training_every
        self._deb_save_before_bod = DEBUG_save_before_bod
        self._deb_alpha = DEBUG_alpha
        self.precision = precision.lower()
        self.rng = rng
        # the default Philox key comes from the OS, not from np.random: the global NumPy
        # stream stays the reference's own (rng='numpy' draws from it)
        self.seed = int.from_bytes(os.urandom(8), 'little') >> 2 if seed is None else int(seed)
        self.write_files = write_files
        self.capacity = capacity
        self._initialize()

    # ------------------------------------------------------------------ set-up
    def _initialize(self):
        """pyvibdmc.py:132-297."""
        T = self.num_timesteps
        self._prop_steps = np.arange(0, T)
        self._branch_step = np.arange(0, T + self.branch_every, self.branch_every)
        self._chkpt_step = np.arange(self.chkpt_every, T + self.chkpt_every, self.chkpt_every)
        self._wfn_save_step = np.arange(self.equil_steps, T + self.wfn_every, self.wfn_every)
        self._desc_wt_save_step = self._wfn_save_step + self.desc_wt_time_steps
        if self._deb_training_every is not None:
            self.deb_train_save_step = np.arange(0, T + self._deb_training_every, self._deb_training_every)
        else:
            self.deb_train_save_step = np.array([], dtype=np.int64)
        self._vref_vs_tau = np.zeros(T)
        self._pop_vs_tau = np.zeros(T)
        if self.start_structures is None:
            raise Exception("Please supply a starting structure for your chemical system.")
        ss = np.asarray(self.start_structures, dtype=np.float64)
        if len(ss.shape) != 3:
            raise Exception("Start structure must have format (n,m,d), where n = 1 or num_walkers, m = num atoms, "
                            "d = dimensions (usually 3)")
        if ss.shape[0] == 1:
            coords = np.repeat(ss, self.num_walkers, axis=0)
        else:
            if ss.shape[0] != self.num_walkers:
                print("WARNING: NUMBER OF STARTING GEOMETRIES DOES NOT EQUAL NUM_WALKERS VARIABLE. MAKE SURE THIS IS"
                      "INTENTIONAL.")
            coords = ss.copy()
        if type(self.atoms) is not list:
            self.atoms = [self.atoms]
        if self.masses is None:
            if '-' in self.atoms[0]:
                self.masses = np.array([Constants.reduced_mass(self.atoms[0])])
                self._atm_nums = get_atomic_num(self.atoms[0].split('-'))
            else:
                self.masses = np.array([Constants.mass(a) for a in self.atoms])
                self._atm_nums = get_atomic_num(self.atoms)
        elif isinstance(self.masses, (int, float)):
            self.masses = np.array([self.masses])
            self._atm_nums = [-1]
        else:
            self.masses = np.array(self.masses, dtype=np.float64)
            self._atm_nums = get_atomic_num(self.atoms)
        if len(self.masses) != len(self.atoms):
            raise Exception("Your number of atoms list does not match your number of masses you provided.")
        if coords.shape[1] != len(self.atoms):
            raise Exception("Your number of atoms list does not match the shape of your walkers.")
        self._sigmas = np.sqrt(self.delta_t / self.masses)
        self._alpha = 1.0 / (2.0 * self.delta_t) if self._deb_alpha is None else self._deb_alpha
        if self.write_files:
            for sub in ('', '/chkpts', '/wfns'):
                os.makedirs(self.output_folder + sub, exist_ok=True)
        self._logger = _Logger(f"{self.output_folder}/{self.sim_name}_log.txt", overwrite=self.cur_timestep == 0,
                               enabled=self.write_files)
        self._thresh_upper = None
        cont_wts = None
        if self.weighting == 'continuous':
            cont_wts = (np.ones(len(coords)) if self.start_cont_wts is None
                        else np.asarray(self.start_cont_wts, dtype=np.float64))
            t = self.cont_wt_thresh
            if t is None:
                self._thresh_lower = 1 / self.num_walkers
            elif isinstance(t, (int, float)):
                self._thresh_lower = t
            elif isinstance(t, list):
                self._thresh_lower = t[0]
                if len(t) == 2:
                    self._thresh_upper = t[1]
            else:
                raise ValueError("Invalid input for continuous weight threshold")
            if self._thresh_upper is not None:
                raise NotImplementedError('the upper continuous-weight threshold is not supported on the GPU path')
        elif self.weighting != 'discrete':
            raise ValueError(f'unknown weighting {self.weighting!r}')
        self._pop_thresh = [self.num_walkers - self.num_walkers * 0.5, self.num_walkers + self.num_walkers * 0.5]
        self.eff_ts = np.zeros(len(self._prop_steps)) if self.impsamp_manager is not None else None
        self._setup_device(coords, cont_wts, self.cur_timestep)
        self._started = False

    def _setup_device(self, coords, cont_wts, step0):
        _torch()
        lib = N.load()
        pes = getattr(self.potential_obj, '_device_pes', None)
        trial = None
        # a trial wave function without a device form (the reference's ImpSampManager with
        # user Python functions) is called on the host around a device move and Metropolis test
        self._htrial = False
        if self.impsamp_manager is not None:
            trial = getattr(self.impsamp_manager, '_device_trial', None)
            if trial is None:
                from .impsamp import ImpSamp
                self._htrial = True
                self.impsamp = ImpSamp(self.impsamp_manager)
                trial = N.TrialDesc()
                trial.kind = N.TRIAL_HOST
        self._dev_pes = pes is not None and pes.kind != N.PES_EXTERNAL
        self._fused = self._dev_pes and not self._htrial
        n, m, d = coords.shape
        np_ = m * d
        cfg = N.Config()
        cfg.num_walkers = int(self.num_walkers)
        cfg.capacity = int(self.capacity)
        cfg.n_atoms = m
        cfg.n_dim = d
        cfg.precision = N.PVD_FP64 if self.precision == 'fp64' else N.PVD_FP32
        cfg.weighting = N.PVD_CONTINUOUS if self.weighting == 'continuous' else N.PVD_DISCRETE
        cfg.branch_every = int(self.branch_every)
        cfg.num_timesteps = int(self.num_timesteps)
        cfg.delta_t = float(self.delta_t)
        cfg.alpha = float(self._alpha)
        cfg.thresh_lower = float(getattr(self, '_thresh_lower', 0.0))
        # sigma and 1/m per component, as the reference builds sigma_trip / inv_masses_trip
        # (pyvibdmc.py:199, 266-274)
        sig = np.repeat(self._sigmas, d)
        invm = (1 / np.repeat(self.masses, d))
        for p in range(np_):
            cfg.sigma[p] = float(sig[p])
            cfg.inv_mass[p] = float(invm[p])
        cfg.seed = self.seed & ((1 << 64) - 1)
        if pes is not None:
            cfg.pes = pes
        else:
            cfg.pes.kind = N.PES_EXTERNAL
        if trial is not None:
            cfg.trial = trial
        self._cfg = cfg
        h = N.ctypes_void()
        N.check(lib.pvd_sim_create(N.byref(cfg), N.byref(h)), 'pvd_sim_create')
        self._h = h
        self._np = np_
        self._shape = (m, d)
        # a surrogate network runs inside the library's step loop (no per-step host work)
        self._surrogate = hasattr(self.potential_obj, 'attach') and pes is not None and pes.kind == N.PES_EXTERNAL
        if self._surrogate:
            self.potential_obj.attach(self)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        w = None if cont_wts is None else np.ascontiguousarray(cont_wts, dtype=np.float64)
        N.check(lib.pvd_sim_set_walkers(self._h, N.ptr(coords), len(coords), N.ptr(w), int(step0),
                                        N.stream_handle()), 'pvd_sim_set_walkers')
        self._n_host = len(coords)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and N._lib is not None:
            try:
                N._lib.pvd_sim_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------ state access
    def _status(self):
        st = N.c_int32()
        fs = N.c_int64()
        n = N.c_int32()
        v = N.c_double()
        nx = N.c_int64()
        rc = N.load().pvd_sim_status(self._h, N.byref(st), N.byref(fs), N.byref(n), N.byref(v), N.byref(nx),
                                     N.stream_handle())
        self._n_host = n.value
        self._vref = v.value
        return rc, st.value, fs.value

    def _raise_if_failed(self):
        rc, st, fs = self._status()
        if rc != 0:
            if fs >= 0:
                self.cur_timestep = int(fs)
            N.check(rc)

    def _get_walkers(self):
        m, d = self._shape
        cap = int(np.ceil(1.5 * self.num_walkers)) + 512
        cap = max(cap, self._n_host + 1)
        buf = np.empty((cap, m, d), dtype=np.float64)
        wb = np.empty(cap, dtype=np.float64) if self.weighting == 'continuous' else None
        n = N.c_int32()
        N.check(N.load().pvd_sim_get_walkers(self._h, N.ptr(buf), N.ptr(wb), cap, N.byref(n), N.stream_handle()),
                'pvd_sim_get_walkers')
        n = n.value
        return buf[:n].copy(), (None if wb is None else wb[:n].copy())

    @property
    def _walker_coords(self):
        return self._get_walkers()[0]

    @property
    def _cont_wts(self):
        return self._get_walkers()[1]

    @property
    def vref_vs_tau(self):
        """Rows [0, cur_timestep) of (step, vref) (pyvibdmc.py:358-363)."""
        v = self._vref_vs_tau[:self.cur_timestep]
        return np.column_stack((np.arange(len(v)), v))

    @property
    def walkers(self):
        """coords, or (coords, weights) for continuous weighting (pyvibdmc.py:365-370)."""
        c, w = self._get_walkers()
        return (c, w) if self.weighting == 'continuous' else c

    def _fetch_traj(self, t0, t1):
        if t1 <= t0:
            return
        cnt = t1 - t0
        v = np.empty(cnt)
        p = np.empty(cnt)
        e = np.empty(cnt) if self.impsamp_manager is not None else None
        N.check(N.load().pvd_sim_trajectory(self._h, t0, cnt, N.ptr(v), N.ptr(p), None, None, N.ptr(e)),
                'pvd_sim_trajectory')
        self._vref_vs_tau[t0:t1] = v
        self._pop_vs_tau[t0:t1] = p
        if e is not None:
            self.eff_ts[t0:t1] = e

    def trajectory_counts(self, t0, t1):
        """births/deaths per step (discrete), or number of splits and the split replay that ran
        (continuous: 0 none; LDS candidate replay: 1 list A, 5 list A cut, 6 lists A and B;
        global-memory candidate replay: 7 list A, 8 after list A alone fell short or the LDS
        form declined its lists; 2-4 interval replay: lists too short / too long / other)."""
        cnt = t1 - t0
        b = np.empty(cnt, dtype=np.int32)
        dd = np.empty(cnt, dtype=np.int32)
        N.check(N.load().pvd_sim_trajectory(self._h, t0, cnt, None, None, N.ptr(b), N.ptr(dd), None))
        return b, dd

    # ------------------------------------------------------------------ stepping
    def _draws(self):
        if self.rng == 'philox':
            return None
        if self.rng == 'numpy':
            return _NumpyDraws
        return self.rng

    def _dev(self, a):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()

    def _host_energy(self, coords):
        v = self.potential(coords)
        return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))

    def _energy_dev(self):
        """V of the current walkers as a device float64 tensor: the device potential in place,
        a device-energy object, or the host potential on a host copy (pyvibdmc.py:786-793)."""
        import torch
        if self._dev_pes:
            v = torch.empty(max(self._n_host, 1), dtype=torch.float64, device='cuda')
            N.check(N.load().pvd_sim_pes(self._h, N.ptr(v), N.stream_handle()), 'pvd_sim_pes')
            return v
        dev_e = getattr(self.potential_obj, 'device_energy', None)
        if dev_e is not None:
            return dev_e(self)
        c, _ = self._get_walkers()
        return self._dev(self._host_energy(c))

    def _drift_to_device(self, cds):
        """ImpSamp.drift on host walkers (imp_samp.py:21-26) -> device (psi, dpsi/psi, d2psi/psi)."""
        f, psi, sd = self.impsamp.drift(cds)
        n = len(cds)
        shape = (n,) + self._shape
        return (self._dev(np.asarray(psi, dtype=np.float64).reshape(n)),
                self._dev(np.asarray(f, dtype=np.float64).reshape(shape)),
                self._dev(np.asarray(sd, dtype=np.float64).reshape(shape)))

    def _start_host_trial(self):
        """First step with a host trial (pyvibdmc.py:760-769, 552-553): V(x0), the local kinetic
        energy of drift(x0), vref; then imp_move_randomly's own drift of the (possibly perturbed)
        walkers fills the caches.  The walkers go back to the device after each drift, since a
        finite-difference trial call perturbs its input in place (imp_samp.py:67-71)."""
        lib = N.load()
        st = N.stream_handle()
        v = self._energy_dev()
        for first in (True, False):
            c, _ = self._get_walkers()
            psi, f, sd = self._drift_to_device(c)
            N.check(lib.pvd_sim_is_cache(self._h, N.ptr(self._dev(c)), N.ptr(psi), N.ptr(f), N.ptr(sd), st),
                    'pvd_sim_is_cache')
            if first:
                N.check(lib.pvd_sim_start_external(self._h, N.ptr(v), st), 'pvd_sim_start_external')
        self._keep = v

    def _host_trial_step(self, t, draws):
        """One importance-sampled step with a host trial (pyvibdmc.py:549-612, 786-848): device
        proposal, the trial and its derivatives at the proposal on the host, device Metropolis,
        potential, device branching."""
        import torch
        lib = N.load()
        st = N.stream_handle()
        n = self._n_host
        m, d = self._shape
        disp = self._dev(draws.displacement(self._sigmas, (n, m, d))) if draws is not None else None
        y = torch.empty(n * m * d, dtype=torch.float64, device='cuda')
        N.check(lib.pvd_sim_is_propose(self._h, N.ptr(disp), N.ptr(y), st), 'pvd_sim_is_propose')
        yh = y.cpu().numpy().reshape(n, m, d)
        psi, f, sd = self._drift_to_device(yh)
        y.copy_(torch.from_numpy(yh.reshape(-1)))
        um = self._dev(draws.uniform(n)) if draws is not None else None
        N.check(lib.pvd_sim_is_accept(self._h, N.ptr(y), N.ptr(psi), N.ptr(f), N.ptr(sd), N.ptr(um), st),
                'pvd_sim_is_accept')
        v = self._energy_dev()
        ub = None
        if draws is not None and self.weighting == 'discrete' and t in self._branch_step:
            ub = self._dev(draws.uniform(n))
        N.check(lib.pvd_sim_branch_external(self._h, N.ptr(v), N.ptr(ub), st), 'pvd_sim_branch_external')
        self._raise_if_failed()

    def _start(self):
        lib = N.load()
        dev_e = getattr(self.potential_obj, 'device_energy', None)
        if self._htrial:
            self._start_host_trial()
        elif self._fused or self._surrogate:
            N.check(lib.pvd_sim_start(self._h, N.stream_handle()), 'pvd_sim_start')
        elif dev_e is not None:
            v = dev_e(self)
            N.check(lib.pvd_sim_start_external(self._h, N.ptr(v), N.stream_handle()), 'pvd_sim_start_external')
        else:
            c, _ = self._get_walkers()
            v = self._dev(self._host_energy(c))
            N.check(lib.pvd_sim_start_external(self._h, N.ptr(v), N.stream_handle()), 'pvd_sim_start_external')
            self._keep = v
        self._started = True

    def _one_step(self, t, draws):
        """One step with host-supplied draws and/or a host potential (pyvibdmc.py:772-848)."""
        lib = N.load()
        n = self._n_host
        m, d = self._shape
        is_ = self.impsamp_manager is not None
        branching = (t in self._branch_step)
        disp = um = ub = None
        if draws is not None:
            disp = self._dev(draws.displacement(self._sigmas, (n, m, d)))
            if is_:
                um = self._dev(draws.uniform(n))
        if self._fused:
            if draws is not None and self.weighting == 'discrete' and branching:
                ub = self._dev(draws.uniform(n))
            if draws is not None:
                N.check(lib.pvd_sim_step_injected(self._h, N.ptr(disp), N.ptr(um), N.ptr(ub), N.stream_handle()),
                        'pvd_sim_step_injected')
            else:
                N.check(lib.pvd_sim_run(self._h, 1, N.stream_handle()), 'pvd_sim_run')
        else:
            N.check(lib.pvd_sim_move(self._h, N.ptr(disp), N.ptr(um), N.stream_handle()), 'pvd_sim_move')
            dev_e = getattr(self.potential_obj, 'device_energy', None)
            if dev_e is not None:
                v = dev_e(self)
            else:
                c, _ = self._get_walkers()
                v = self._dev(self._host_energy(c))
            if draws is not None and self.weighting == 'discrete' and branching:
                ub = self._dev(draws.uniform(n))
            N.check(lib.pvd_sim_branch_external(self._h, N.ptr(v), N.ptr(ub), N.stream_handle()),
                    'pvd_sim_branch_external')
        self._raise_if_failed()

    def _run_chunk(self, t0, t1):
        draws = self._draws()
        dev_e = getattr(self.potential_obj, 'device_energy', None)
        lib = N.load()
        if self._htrial:
            for t in range(t0, t1):
                self._host_trial_step(t, draws)
        elif draws is None and (self._fused or self._surrogate):
            N.check(lib.pvd_sim_run(self._h, t1 - t0, N.stream_handle()), 'pvd_sim_run')
            self._raise_if_failed()
        elif draws is None and dev_e is not None:
            # another device potential outside the fused kernel: move, V on the device,
            # branch - no host synchronisation inside the chunk
            st = N.stream_handle()
            for t in range(t0, t1):
                N.check(lib.pvd_sim_move(self._h, None, None, st), 'pvd_sim_move')
                v = dev_e(self)
                N.check(lib.pvd_sim_branch_external(self._h, N.ptr(v), None, st), 'pvd_sim_branch_external')
            self._raise_if_failed()
        else:
            for t in range(t0, t1):
                self._one_step(t, draws)
        self._fetch_traj(t0, t1)

    # ------------------------------------------------------------------ the loop
    def propagate(self):
        """pyvibdmc.py:701-876, with the steps between host events fused on the device."""
        steps = self._prop_steps
        if len(steps) == 0:
            return
        first, T = int(steps[0]), int(steps[-1]) + 1
        bounds = set(int(x) for x in self._chkpt_step) | set(int(x) for x in self._wfn_save_step) \
            | set(int(x) for x in self._desc_wt_save_step) | set(int(x) + 1 for x in self.deb_train_save_step)
        bounds = sorted(b for b in bounds if first < b < T) + [T]
        t = first
        if self.weighting == 'discrete':
            self._status()
            if self._n_host < self._pop_thresh[0] or self._n_host > self._pop_thresh[1]:
                raise N.DMCError(N.ST_POP, N.MASSIVE)
        self._logger.write(f"Starting at time step {first}\n")
        for b in bounds:
            self.cur_timestep = t
            if t in self._chkpt_step:
                self._logger.write(f"Checkpointing, time step {t}\n")
                self._delete_older_checkpoints(t)
                self.chkpt(t)
            if t in self._wfn_save_step:
                self._logger.write(f"Starting descendant weighting, time step {t}\n")
                N.check(N.load().pvd_sim_desc_begin(self._h, N.stream_handle()), 'pvd_sim_desc_begin')
            if not self._started:
                self._start()
            self._run_chunk(t, b)
            self.cur_timestep = b - 1
            if b in self._desc_wt_save_step:
                self._logger.write(f"Finished descendant weighting at end of time step {b - 1}\n")
                self._save_wfn(b - self.desc_wt_time_steps)
            if (b - 1) in self.deb_train_save_step:
                self._save_training(b - 1)
            t = b

    def _save_wfn(self, start_step):
        m, d = self._shape
        cap = int(np.ceil(1.5 * self.num_walkers)) + 512
        desc = np.empty(cap)
        par = np.empty((cap, m, d))
        pw = np.empty(cap) if self.weighting == 'continuous' else None
        npar = N.c_int32()
        N.check(N.load().pvd_sim_desc_end(self._h, N.ptr(desc), N.ptr(par), N.ptr(pw), N.byref(npar),
                                          N.stream_handle()), 'pvd_sim_desc_end')
        k = npar.value
        rec = dict(coords=par[:k].copy(), desc_wts=desc[:k].copy())
        if pw is not None:
            rec['parent_wts'] = pw[:k].copy()
        self.wfns = getattr(self, 'wfns', [])
        self.wfns.append((start_step, rec))
        if self.write_files:
            # SimArchivist.save_h5 of pyvibdmc.py:861-869
            hdf5.write(f"{self.output_folder}/wfns/{self.sim_name}_wfn_{start_step}ts.hdf5", list(rec), list(rec.values()))

    def _save_training(self, step):
        c, _ = self._get_walkers()
        pots = np.asarray(self.potential(c), dtype=np.float64)
        if self.write_files:
            # pyvibdmc.py:800-804, 840-844
            hdf5.write(f"{self.output_folder}/{self.sim_name}_training_{step}ts.hdf5", ['coords', 'pots'], [c, pots])
        self.training = getattr(self, 'training', [])
        self.training.append((step, c, pots))

    def run(self):
        """pyvibdmc.py:878-932: propagate, always leave a final checkpoint, report the ZPE."""
        print("Starting Simulation...")
        start = time.time()
        err = None
        try:
            self.propagate()
            self._delete_older_checkpoints(self.cur_timestep)
        except Exception as e:
            import traceback
            print("ERROR! An error occurred while running the DMC simulation. Dumping a final checkpoint...")
            print("Ignore Approximate ZPE!!!")
            traceback.print_exc()
            err = e
        finally:
            try:
                self.chkpt(self.cur_timestep)
            except Exception:
                pass
            vref_cm = Constants.convert(self._vref_vs_tau, "wavenumbers", to_AU=False)
            print("Simulation Complete")
            print('Approximate ZPE', np.average(vref_cm[len(vref_cm) // 4:]))
            ts = self.eff_ts if self.impsamp_manager is not None else np.arange(len(vref_cm)) * self.delta_t
            self.sim_info = dict(vref_vs_tau=np.column_stack((ts, self._vref_vs_tau)),
                                 pop_vs_tau=np.column_stack((ts, self._pop_vs_tau)),
                                 atomic_nums=np.asarray(self._atm_nums), atomic_masses=np.asarray(self.masses))
            if self.write_files:
                # pyvibdmc.py:918-922
                hdf5.write(f"{self.output_folder}/{self.sim_name}_sim_info.hdf5", list(self.sim_info),
                           list(self.sim_info.values()))
            self.run_time = time.time() - start
            self._logger.write(f"Simulation took {self.run_time} seconds.\n\n")
            self._logger.close()
        if err is not None:
            raise err

    def zpe(self, onwards=None, cm=True):
        """Mean vref from ``onwards`` (default: the last three quarters)."""
        v = self._vref_vs_tau
        o = len(v) // 4 if onwards is None else onwards
        z = np.average(v[o:])
        return Constants.convert(z, "wavenumbers", to_AU=False) if cm else z

    # ------------------------------------------------------------------ checkpoints
    def _delete_older_checkpoints(self, step):
        if not self.write_files:
            return
        import glob
        for p in glob.glob(f'{self.output_folder}/chkpts/{self.sim_name}*.pickle'):
            try:
                if int(p.split('_')[-1].split('.')[0]) < step:
                    os.remove(p)
            except ValueError:
                pass

    def _state(self):
        c, w = self._get_walkers()
        keep = {k: v for k, v in self.__dict__.items()
                if k not in ('potential', 'potential_obj', 'potential_info', 'impsamp_manager', 'impsamp', '_h',
                             '_cfg', '_logger', '_keep', 'rng')}
        keep = copy.deepcopy(keep)
        keep['_saved_coords'] = c
        keep['_saved_wts'] = w
        keep['_rng_kind'] = self.rng if isinstance(self.rng, str) else 'numpy'
        return keep

    def chkpt(self, step):
        """Pickle of the host-side state with the walkers (and weights) copied out
        (sim_archive.py:26-32).  The Philox key and step are part of it, so a restart
        continues the same draw sequence."""
        if not self.write_files:
            return
        state = self._state()
        with open(f'{self.output_folder}/chkpts/{self.sim_name}_{step}.pickle', 'wb') as fh:
            pickle.dump(state, fh, protocol=4)


def dmc_restart(potential, chkpt_folder, sim_name, additional_timesteps=0, impsamp=None, imp_samp_oned=False,
                fixed_node=None):
    """Reload a checkpoint written by ``DMC_Sim.chkpt`` and extend the run (pyvibdmc.py:949-965)."""
    import glob
    files = sorted(glob.glob(f"{chkpt_folder}/chkpts/{sim_name}_*.pickle"),
                   key=lambda p: int(p.split('_')[-1].split('.')[0]))
    if not files:
        raise FileNotFoundError(f'no checkpoint for {sim_name} in {chkpt_folder}/chkpts')
    with open(files[-1], 'rb') as fh:
        st = pickle.load(fh)
    sim = DMC_Sim.__new__(DMC_Sim)
    sim.__dict__.update({k: v for k, v in st.items() if not k.startswith('_saved')})
    add = int(additional_timesteps)
    sim.num_timesteps = sim.num_timesteps + add
    T = sim.num_timesteps
    sim._vref_vs_tau = np.concatenate((sim._vref_vs_tau, np.zeros(add)))
    sim._pop_vs_tau = np.concatenate((sim._pop_vs_tau, np.zeros(add)))
    sim._branch_step = np.arange(0, T + sim.branch_every, sim.branch_every)
    sim._chkpt_step = np.arange(sim.chkpt_every, T + sim.chkpt_every, sim.chkpt_every)
    sim._wfn_save_step = np.arange(sim.equil_steps, T + sim.wfn_every, sim.wfn_every)
    sim._desc_wt_save_step = sim._wfn_save_step + sim.desc_wt_time_steps
    sim._prop_steps = np.arange(sim.cur_timestep, T)
    sim.potential_obj = potential
    sim.potential = potential.getpot
    sim.potential_info = dict(vars(potential))
    sim.impsamp_manager = impsamp
    sim.imp1d = imp_samp_oned
    sim.rng = st.get('_rng_kind', 'philox')
    if impsamp is not None:
        sim.eff_ts = np.concatenate((sim.eff_ts if sim.eff_ts is not None else np.zeros(T - add), np.zeros(add)))
    sim._logger = _Logger(f"{sim.output_folder}/{sim.sim_name}_log.txt", enabled=sim.write_files)
    sim._setup_device(st['_saved_coords'], st['_saved_wts'], sim.cur_timestep)
    sim._started = False
    return sim
