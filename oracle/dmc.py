"""FP64 NumPy restatement of the reference DMC step loop, with injectable draws.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``pyvibdmc/pyvibdmc.py`` (reference) statement by statement for the parts on
the propagation path; I/O (logger, checkpoints, HDF5) is left out because it does not
change any number.  Cited lines are in the reference file unless stated otherwise.

Pinning: ``tests/test_oracle_golden.py`` runs this class against trajectories written by
the reference itself (``tests/golden/gen_golden.py``) in discrete, continuous,
descendant-weighting and importance-sampling modes and requires bit-identical
populations, survivor indices and coordinates and vref within 1e-12 hartree.  The
discrete 1-D oscillator ground-state energy is further pinned by ``oracle/eigen.py``.
"""
import numpy as np


class LegacyDraws:
    """The reference's draw stream: global legacy ``np.random`` (MT19937), seeded by the caller.

    * displacement: ``np.random.normal(0.0, sigmas, size=(N, d, M)).transpose(0, 2, 1)``
      (pyvibdmc.py:544-546, 556-558);
    * uniforms: ``np.random.random(N)`` (pyvibdmc.py:392, 601).
    ``RandomState(seed)`` reproduces ``np.random.seed(seed)`` + global calls exactly."""

    def __init__(self, seed):
        self.rs = np.random.RandomState(seed)
        self.log = None          # optional list collecting every draw, for feeding the GPU

    def displacement(self, sigmas, shape):
        n, m, d = shape
        disp = self.rs.normal(0.0, sigmas, size=(n, d, m)).transpose(0, 2, 1)
        if self.log is not None:
            self.log.append(('disp', np.ascontiguousarray(disp)))
        return disp

    def uniform(self, n):
        u = self.rs.random_sample(n)
        if self.log is not None:
            self.log.append(('u', u.copy()))
        return u


class MassiveEvent(ValueError):
    pass


MASSIVE = "Massive walker birth or death event!!!!!!! Dying..."


class OracleDMC:
    """State machine of ``DMC_Sim`` (pyvibdmc.py:18-932) on the propagation path.

    ``potential``: callable (N,M,d) -> (N,) hartree.
    ``trial``: object with ``trial(cds)`` and ``derivs(cds) -> (dpsi/psi, d2psi/psi)``
    (the ImpSampManager contract, imp_samp.py:16-26), or None.
    """

    def __init__(self, *, potential, masses, start_structures, num_walkers, num_timesteps,
                 delta_t=1, weighting='discrete', equil_steps=2000, wfn_every=1000,
                 desc_wt_steps=100, branch_every=1, cont_wt_thresh=None, start_cont_wts=None,
                 trial=None, imp_samp_oned=False, DEBUG_alpha=None, draws=None, record_idx=False):
        self.potential = potential
        self.num_walkers = num_walkers
        self.num_timesteps = num_timesteps
        self.delta_t = delta_t
        self.weighting = weighting.lower()
        self.branch_every = branch_every
        self.draws = draws
        self.record_idx = record_idx
        # schedules, pyvibdmc.py:138-148
        self._prop_steps = np.arange(0, num_timesteps)
        self._branch_step = np.arange(0, num_timesteps + branch_every, branch_every)
        self._wfn_save_step = np.arange(equil_steps, num_timesteps + wfn_every, wfn_every)
        self._desc_wt_save_step = self._wfn_save_step + desc_wt_steps
        self.desc_wt_time_steps = desc_wt_steps
        self._vref_vs_tau = np.zeros(num_timesteps)
        self._pop_vs_tau = np.zeros(num_timesteps)
        # walkers, pyvibdmc.py:156-169
        ss = np.asarray(start_structures, dtype=np.float64)
        if ss.shape[0] == 1:
            self.coords = np.repeat(ss, num_walkers, axis=0)
        else:
            self.coords = ss.copy()
        self.masses = np.atleast_1d(np.asarray(masses, dtype=np.float64))
        # sigma and alpha, pyvibdmc.py:199-203
        self._sigmas = np.sqrt(self.delta_t / self.masses)
        self._alpha = 1.0 / (2.0 * self.delta_t) if DEBUG_alpha is None else DEBUG_alpha
        # continuous weights, pyvibdmc.py:214-233
        self._thresh_upper = None
        if self.weighting == 'continuous':
            self.cont_wts = (np.ones(num_walkers) if start_cont_wts is None
                             else np.array(start_cont_wts, dtype=np.float64))
            if cont_wt_thresh is None:
                self._thresh_lower = 1 / num_walkers
            elif isinstance(cont_wt_thresh, (int, float)):
                self._thresh_lower = cont_wt_thresh
            elif isinstance(cont_wt_thresh, list):
                self._thresh_lower = cont_wt_thresh[0]
                if len(cont_wt_thresh) == 2:
                    self._thresh_upper = cont_wt_thresh[1]
            else:
                raise ValueError("Invalid input for continuous weight threshold")
        else:
            self.cont_wts = None
        self._desc_wt = False
        self._who_from = None
        self._pop_thresh = [num_walkers - num_walkers * 0.5, num_walkers + num_walkers * 0.5]
        # importance sampling, pyvibdmc.py:257-274
        self.trial = trial
        self.f_x = self.psi_1 = self.psi_sec_der = None
        if trial is not None:
            m = self.masses
            self.inv_masses_trip = (1 / np.repeat(m, 3)).reshape(len(m), 3)[np.newaxis, ...]
            self.sigma_trip = np.repeat(self._sigmas, 3).reshape(len(m), 3)[np.newaxis, ...]
            self.eff_ts = np.zeros(len(self._prop_steps))
            if imp_samp_oned:
                self.sigma_trip = self._sigmas
                self.inv_masses_trip = (1 / m)[np.newaxis]
        self.V = None
        self.vref = None
        self.cur_timestep = 0
        self._next = 0           # first step of the next propagate() chunk
        # per-step records
        self.rec = dict(pop_in=[], idx=[], births=[], deaths=[], accept=[], n_branched=[], src=[])
        self.wfns = []          # (start step, parent coords, desc_wts, parent_wts)

    # ---- IS helpers, imp_samp.py:21-53 -------------------------------------------------
    def drift(self, cds):
        psi = self.trial.trial(cds)
        d, s = self.trial.derivs(cds)
        return d, psi, s

    @staticmethod
    def metropolis(sigma_trip, trial_x, trial_y, disp_x, disp_y, D_x, D_y, dt):
        """imp_samp.py:28-47."""
        psi_ratio = (trial_y / trial_x) ** 2
        term_1 = np.exp(-1 * (disp_x - disp_y - D_y * dt) ** 2 / (2 * sigma_trip ** 2))
        term_2 = np.exp(-1 * (disp_y - disp_x - D_x * dt) ** 2 / (2 * sigma_trip ** 2))
        accep = term_1 / term_2
        if accep.shape[-1] == 1:
            accep = accep.squeeze() * psi_ratio.squeeze()
        else:
            accep = np.prod(np.prod(accep, axis=1), axis=1) * psi_ratio
        flipped = np.where(trial_x * trial_y <= 0)[0]
        accep[flipped] = 0.0
        return accep

    @staticmethod
    def local_kin(inv_masses_trip, sec_deriv):
        """imp_samp.py:49-53."""
        return -0.5 * np.sum(np.sum(inv_masses_trip * sec_deriv, axis=1), axis=1)

    # ---- moves ---------------------------------------------------------------------------
    def move_randomly(self):
        """pyvibdmc.py:540-547."""
        self.coords += self.draws.displacement(self._sigmas, self.coords.shape)

    def imp_move_randomly(self):
        """pyvibdmc.py:549-612 (excited-state branch excluded)."""
        if self.f_x is None or self.psi_1 is None:
            self.f_x, self.psi_1, self.psi_sec_der = self.drift(self.coords)
        disps = self.draws.displacement(self._sigmas, self.coords.shape)
        d_x = self.inv_masses_trip * self.f_x
        displaced = self.coords + disps + d_x * self.delta_t
        f_y, psi_2, sd_y = self.drift(displaced)
        d_y = self.inv_masses_trip * f_y
        met = self.metropolis(self.sigma_trip, self.psi_1, psi_2, self.coords, displaced, d_x, d_y,
                              self.delta_t)
        randos = self.draws.uniform(len(self.coords))
        accept = np.argwhere(met > randos)
        self.dt_factor = len(accept) / len(self.coords)
        self.coords[accept] = displaced[accept]
        self.f_x[accept] = f_y[accept]
        self.psi_1[accept] = psi_2[accept]
        self.psi_sec_der[accept] = sd_y[accept]
        self.rec['accept'].append(len(accept))
        return len(self.coords) - len(accept)

    def update_effective_timestep(self):
        """pyvibdmc.py:372-378."""
        dt = self.delta_t * self.dt_factor
        t = self.cur_timestep
        self.eff_ts[t] = dt if t == 0 else self.eff_ts[t - 1] + dt
        return dt

    # ---- branching -----------------------------------------------------------------------
    def _branch(self, walkers_below):
        """pyvibdmc.py:340-356."""
        for walker in walkers_below:
            mx = np.argmax(self.cont_wts)
            self.coords[walker] = np.copy(self.coords[mx])
            self.V[walker] = np.copy(self.V[mx])
            if self._desc_wt:
                self._who_from[walker] = self._who_from[mx]
            self.cont_wts[mx] /= 2.0
            self.cont_wts[walker] = np.copy(self.cont_wts[mx])
            if self.trial is not None:
                self.f_x[walker] = np.copy(self.f_x[mx])
                self.psi_1[walker] = np.copy(self.psi_1[mx])
                self.psi_sec_der[walker] = np.copy(self.psi_sec_der[mx])

    def birth_or_death(self):
        """pyvibdmc.py:380-454."""
        dt = self.delta_t
        if self.trial is not None:
            dt = self.update_effective_timestep()
        if self.weighting == 'discrete':
            rand_nums = self.draws.uniform(len(self.coords))
            weights = np.exp(-1. * (self.V - self.vref) * dt)
            if not np.all(np.isfinite(weights)) or np.any(weights > self._pop_thresh[1] + 1):
                raise MassiveEvent(MASSIVE)
            counts = np.floor(weights).astype(np.int64)
            counts += rand_nums < (weights - counts)
            num_deaths = int(np.count_nonzero(counts == 0))
            num_births = int(np.sum(np.maximum(counts - 1, 0)))
            final_pop = int(np.sum(counts))
            if final_pop < self._pop_thresh[0] or final_pop > self._pop_thresh[1]:
                raise MassiveEvent(MASSIVE)
            idx = np.repeat(np.arange(len(counts)), counts)
            self.coords = self.coords[idx]
            self.V = self.V[idx]
            if self.trial is not None:
                self.f_x = self.f_x[idx]
                self.psi_1 = self.psi_1[idx]
                self.psi_sec_der = self.psi_sec_der[idx]
            if self._desc_wt:
                self._who_from = self._who_from[idx]
            if self.record_idx:
                self.rec['idx'].append(idx.astype(np.int32))
            self.rec['births'].append(num_births)
            self.rec['deaths'].append(num_deaths)
            return num_births, num_deaths, final_pop
        else:
            self.cont_wts *= np.exp(-1.0 * (self.V - self.vref) * dt)
            kill = np.where(self.cont_wts < self._thresh_lower)[0]
            if self.record_idx:
                # resolved source of every slot after the sequential split
                saved = (self._desc_wt, self._who_from)
                self._desc_wt, self._who_from = True, np.arange(len(self.coords))
                self._branch(kill)
                src = self._who_from
                self._desc_wt, self._who_from = saved
                if self._desc_wt:
                    self._who_from = self._who_from[src]
                self.rec['src'].append(src.astype(np.int32))
            else:
                self._branch(kill)
            if self._thresh_upper is not None:
                n_above = np.sum(self.cont_wts > self._thresh_upper)
                kill_up = np.argpartition(self.cont_wts, n_above)[:n_above]
                self._branch(kill_up)
            else:
                kill_up = []
            nb = len(kill) + len(kill_up)
            self.rec['n_branched'].append(nb)
            return nb, np.amax(self.cont_wts), np.amin(self.cont_wts)

    def calc_vref(self):
        """pyvibdmc.py:651-661."""
        if self.weighting == 'discrete':
            v_bar = np.average(self.V)
            corr = (len(self.V) - self.num_walkers) / self.num_walkers
        else:
            v_bar = np.average(self.V, weights=self.cont_wts)
            corr = (np.sum(self.cont_wts) - self.num_walkers) / self.num_walkers
        self.vref = v_bar - (self._alpha * corr)

    def calc_desc_wts(self):
        """pyvibdmc.py:663-672 (continuous branch via an equivalent grouped sum)."""
        if self.weighting == 'discrete':
            unique, counts = np.unique(self._who_from, return_counts=True)
            self._desc_wts[unique] = counts
        else:
            n = len(self.cont_wts)
            self._desc_wts[:n] = np.bincount(self._who_from, weights=self.cont_wts, minlength=n)[:n]

    # ---- the loop, pyvibdmc.py:701-876 -----------------------------------------------------
    def propagate(self, n_steps=None):
        hi = self.num_timesteps if n_steps is None else self._next + n_steps
        steps = self._prop_steps[(self._prop_steps >= self._next) & (self._prop_steps < hi)]
        for t in steps:
            self.cur_timestep = t
            if self.weighting == 'discrete':
                cur = len(self.coords)
                if cur < self._pop_thresh[0] or cur > self._pop_thresh[1]:
                    raise MassiveEvent(MASSIVE)
            self.rec['pop_in'].append(len(self.coords))
            if t in self._wfn_save_step:
                self._desc_wts = np.zeros(len(self.coords))
                self._parent = np.copy(self.coords)
                self._parent_wts = None if self.cont_wts is None else np.copy(self.cont_wts)
                self._who_from = np.arange(len(self.coords))
                self._desc_wt = True
            if t == self._prop_steps[0]:
                self.V = self.potential(self.coords)
                if self.trial is not None:
                    _, _, sd = self.drift(self.coords)
                    self.psi_sec_der = sd
                    self.V = self.V + self.local_kin(self.inv_masses_trip, self.psi_sec_der)
                self.calc_vref()
                self.vref0 = self.vref
            if self.trial is None:
                self.move_randomly()
            else:
                self.imp_move_randomly()
            self.V = self.potential(self.coords)
            if self.trial is not None:
                self.V = self.V + self.local_kin(self.inv_masses_trip, self.psi_sec_der)
            if t in self._branch_step:
                self.birth_or_death()
            elif self.weighting == 'continuous':
                dt = self.delta_t
                if self.trial is not None:
                    dt = self.update_effective_timestep()
                self.cont_wts *= np.exp(-1.0 * (self.V - self.vref) * dt)
            self.calc_vref()
            self._vref_vs_tau[t] = self.vref
            self._pop_vs_tau[t] = len(self.coords) if self.weighting == 'discrete' else np.sum(self.cont_wts)
            if t + 1 in self._desc_wt_save_step:
                self._desc_wt = False
                self.calc_desc_wts()
                self.wfns.append((int(t + 1 - self.desc_wt_time_steps), self._parent.copy(),
                                  self._desc_wts.copy(),
                                  None if self._parent_wts is None else self._parent_wts.copy()))
        if len(steps):
            self._next = int(steps[-1]) + 1
        return self

    @property
    def vref_vs_tau(self):
        """pyvibdmc.py:358-363: rows [0, cur_timestep) where cur_timestep is the last step run."""
        v = self._vref_vs_tau[:self.cur_timestep]
        return np.column_stack((np.arange(len(v)), v))

    def zpe_cm(self):
        """Mean of vref over the last 3/4 of the run in cm^-1 (pyvibdmc.py:904-908)."""
        from .pes import WAVENUMBERS
        v = self._vref_vs_tau / WAVENUMBERS
        return np.average(v[len(v) // 4:])
