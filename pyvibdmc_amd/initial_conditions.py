"""Initial ensembles: harmonic analysis of a potential and normal-mode sampling.

Mirrors simulation_utilities/initial_conditioner/ (``HarmonicAnalysis``
harmonic_analysis.py:11-129, ``InitialConditioner`` initial_conditioner.py:7-108,
``MolFiniteDifference`` finite_difference.py:6-68) with the same stencils, weights and
order of operations.  The Hessian's finite-difference stencils (9 geometries per
off-diagonal element, 5 per diagonal one) are evaluated in ONE ``getpot`` call over the
whole stack, so a device potential computes the full Hessian in a single kernel launch
instead of one call per element; the derivatives are formed from those energies exactly
as the reference forms them.
"""
import itertools as itt

import numpy as np
import numpy.linalg as la

from .constants import Constants

__all__ = ['MolFiniteDifference', 'HarmonicAnalysis', 'InitialConditioner']


class MolFiniteDifference:
    """Cartesian finite differences of a molecule (finite_difference.py:6-68)."""
    weights_der1 = {'3': np.array([-1 / 2, 0, 1 / 2]), '5': np.array([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12])}
    weights_der2 = {'3': np.array([1, -2, 1]), '5': np.array([-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])}

    @staticmethod
    def displace_molecule(eq_geom, atm_cd, dx, num_disps):
        """Stencil geometries: one (atom, xyz) displaced over ``num_disps`` points, or two
        (atom, xyz) pairs displaced over the product grid (first index slowest)."""
        half = num_disps // 2
        steps = np.arange(-half, half + 1) * dx
        if len(atm_cd) == 2:
            out = np.zeros((len(steps),) + np.shape(eq_geom))
            for k, s in enumerate(steps):
                d = np.zeros(eq_geom.shape)
                d[atm_cd[0], atm_cd[1]] += s
                out[k] = eq_geom + d
            return out
        a1, c1, a2, c2 = atm_cd
        out = np.zeros((len(steps) ** 2,) + np.shape(eq_geom))
        k = 0
        for s1 in steps:
            for s2 in steps:
                d = np.zeros(eq_geom.shape)
                d[a1, c1] += s1
                d[a2, c2] += s2
                out[k] = eq_geom + d
                k += 1
        return out

    @classmethod
    def differentiate(cls, values, dx, num_points, der):
        if der == 1:
            return np.dot(cls.weights_der1[str(num_points)] / dx, values)
        if der == 11:
            w = cls.weights_der1[str(num_points)] / dx
            return np.dot(w, np.dot(w, values))
        if der == 2:
            return np.dot(cls.weights_der2[str(num_points)] / dx ** 2, values)
        raise ValueError(f'unknown derivative {der}')


class HarmonicAnalysis:
    """Mass-weighted Cartesian Hessian by finite differences of ``potential.getpot`` and
    its eigen-decomposition (harmonic_analysis.py:11-129)."""

    def __init__(self, eq_geom, atoms, potential, masses=None, dx=1.0e-3, points_diag=5, points_off_diag=3,
                 dipole=None):
        self.eq_geom = np.asarray(eq_geom, dtype=np.float64)
        self.atoms = atoms
        self.potential = potential
        self.masses = masses
        self.dx = dx
        self.points_diag = points_diag
        self.points_off_diag = points_off_diag
        self.num_elems = 3 * len(atoms)
        self.dipole = dipole
        if self.masses is None:
            self.masses = np.array([Constants.mass(a) for a in self.atoms])
        elif isinstance(self.masses, list):
            self.masses = np.array(self.masses)

    def _stencils(self):
        """Off-diagonal (i < j, row order) then diagonal stencils, as the reference visits them."""
        m = len(self.atoms)
        n = self.num_elems
        flat = np.ravel_multi_index(np.array(np.triu_indices(n, 1)), [n, n])
        off_info = np.array(np.unravel_index(flat, (m, 3, m, 3))).T
        off = [MolFiniteDifference.displace_molecule(self.eq_geom, off_info[k], self.dx, self.points_off_diag)
               for k in range(len(off_info))]
        diag_info = list(itt.product(np.arange(m), np.arange(3)))
        diag = [MolFiniteDifference.displace_molecule(self.eq_geom, diag_info[k], self.dx, self.points_diag)
                for k in range(n)]
        return off, diag

    def generate_hessian(self):
        n = self.num_elems
        off, diag = self._stencils()
        stack = np.concatenate(off + diag)
        v = np.asarray(self.potential.getpot(stack), dtype=np.float64).reshape(-1)  # one batched call
        hess = np.zeros((n, n))
        pairs = list(itt.combinations(np.arange(n), 2))
        o, k9 = 0, self.points_off_diag ** 2
        for k in range(len(pairs)):
            potz = np.reshape(v[o:o + k9], (self.points_off_diag, self.points_off_diag))
            hess[pairs[k]] = MolFiniteDifference.differentiate(potz, self.dx, self.points_off_diag, 11)
            o += k9
        hess = hess + hess.T
        for k in range(n):
            hess[k, k] = MolFiniteDifference.differentiate(v[o:o + self.points_diag], self.dx, self.points_diag, 2)
            o += self.points_diag
        return hess

    def dipole_derivs(self):
        """d mu / d x for each Cartesian coordinate from a dipole surface's ``getpot``
        returning (n, 3) (harmonic_analysis.py:100-118)."""
        m = len(self.atoms)
        info = list(itt.product(np.arange(m), np.arange(3)))
        stack = np.concatenate([MolFiniteDifference.displace_molecule(self.eq_geom, info[k], self.dx, 5)
                                for k in range(self.num_elems)])
        dip = np.asarray(self.dipole.getpot(stack)).reshape(self.num_elems, 5, 3)
        out = np.zeros((self.num_elems, 3))
        for k in range(self.num_elems):
            for i in range(3):
                out[k, i] = MolFiniteDifference.differentiate(dip[k, :, i], self.dx, 5, 1)
        return out

    def diagonalize(self, hessian):
        md = np.repeat(self.masses, 3)
        hess_mw = hessian / np.sqrt(md) / np.sqrt(md[:, np.newaxis])
        freqs, modes = la.eigh(hess_mw)
        freqs_cm = Constants.convert(np.sqrt(np.abs(freqs)) * np.sign(freqs), 'wavenumbers', to_AU=False)
        return freqs_cm, modes

    def run(self):
        return self.diagonalize(self.generate_hessian())


class InitialConditioner:
    """Starting ensembles (initial_conditioner.py:7-108): ``harmonic_sampling`` draws the
    harmonic ground state along the 3N-6 normal modes; ``permute_atoms`` shuffles like atoms.
    Draws come from the global ``np.random`` in the reference's order."""

    def __init__(self, coord, atoms, num_walkers, technique, technique_kwargs, masses=None):
        self.coord = coord
        self.atoms = atoms
        self.masses = masses
        self.num_walkers = num_walkers
        self.technique = technique.lower()
        self.technique_kwargs = technique_kwargs
        if self.technique == 'harmonic_sampling':
            self.run_func = self.run_harm
        elif self.technique == 'permute_atoms':
            self.run_func = self.run_permute
        else:
            raise NotImplementedError("Choose another initial condition: harmonic_sampling or permute_atoms")

    def gen_disps(self, sigmas):
        return np.random.normal(loc=0, scale=sigmas, size=(self.num_walkers, len(sigmas)))

    def displace_along_nms(self, freqz, nmz, massez, ensemble):
        three_n = len(freqz) + 6
        disp_nm = self.gen_disps(1 / np.sqrt(freqz))
        carts = np.matmul(disp_nm, nmz.T[np.newaxis, :, :]).squeeze()
        carts = np.reshape(carts, (-1, three_n // 3, 3)) / np.sqrt(massez)[:, np.newaxis]
        base = np.tile(self.coord, (self.num_walkers, 1, 1)) if ensemble is None else ensemble
        return base + carts

    def run_harm(self):
        kw = self.technique_kwargs
        freqs, nms, scaling, ensemble = kw['freqs'], kw['normal_modes'], kw['scaling_factor'], kw['ensemble']
        masses = np.array([Constants.mass(a) for a in self.atoms]) if self.masses is None else self.masses
        freqs_au = Constants.convert(freqs[6:], 'wavenumbers', to_AU=True) / scaling
        return self.displace_along_nms(freqs_au, nms[:, 6:], masses, ensemble)

    def run_permute(self):
        walkers = (np.tile(self.coord, (self.num_walkers, 1, 1)) if self.technique_kwargs['ensemble'] is None
                   else self.technique_kwargs['ensemble'])
        for group in self.technique_kwargs['like_atoms']:
            sub = walkers[:, group]
            [np.random.shuffle(x) for x in sub]
            walkers[:, group] = sub
        return walkers

    def run(self):
        return self.run_func()
