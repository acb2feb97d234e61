"""SimInfo: the reference's reader of simulation results (analysis/extract_sim_info.py:9-133),
over the h5py-free HDF5 reader.  Reads the reference's own files as well as this package's."""
import numpy as np

from . import hdf5
from .constants import Constants

__all__ = ['SimInfo']


class SimInfo:
    """Results of one simulation: ``{sim_name}_sim_info.hdf5`` and its ``wfns/`` directory."""

    def __init__(self, h5_name):
        self.fname = h5_name
        self._initialize()
        self._load_sim_H5()

    def _initialize(self):
        parts = self.fname.split('/')
        sim_name = parts[-1].split('sim_info')[0]
        pth = '/'.join(parts[:-1])
        self._wfn_names = f"{pth}/wfns/{sim_name}wfn_"

    def _load_sim_H5(self):
        d = hdf5.read(self.fname)
        self.vref_vs_tau = d['vref_vs_tau']
        self.pop_vs_tau = d['pop_vs_tau']
        self.atom_nums = d['atomic_nums']
        self.atom_masses = d['atomic_masses']

    @staticmethod
    def get_wfn(wfn_fl, ret_ang=False, get_parent_wts=False):
        """Walkers (n, m, 3) and descendant weights of one wave-function file."""
        d = hdf5.read(wfn_fl)
        cds = d['coords']
        if ret_ang:
            cds = Constants.convert(cds, 'angstroms', to_AU=False)
        if get_parent_wts:
            return cds, d['desc_wts'], d['parent_wts']
        return cds, d['desc_wts']

    def get_wfns(self, time_step_list, ret_ang=False, get_parent_wts=False):
        """Wave functions of the given time steps, concatenated."""
        steps = [time_step_list] if isinstance(time_step_list, (int, np.integer)) else time_step_list
        got = [self.get_wfn(f'{self._wfn_names}{x}ts.hdf5', ret_ang, get_parent_wts) for x in steps]
        cds = np.concatenate([g[0] for g in got])
        dw = np.concatenate([g[1] for g in got])
        if get_parent_wts:
            return cds, dw, np.concatenate([g[2] for g in got])
        return cds, dw

    def get_vref(self, ret_cm=False):
        if ret_cm:
            vref_cm = Constants.convert(self.vref_vs_tau[:, 1], 'wavenumbers', to_AU=False)
            return np.column_stack((self.vref_vs_tau[:, 0], vref_cm))
        return self.vref_vs_tau

    def get_pop(self):
        return self.pop_vs_tau

    def get_atomic_nums(self):
        return self.atom_nums

    def get_atom_masses(self):
        return self.atom_masses

    def get_zpe(self, onwards=1000, ret_cm=False):
        """Mean vref from index ``onwards`` (extract_sim_info.py:105-111)."""
        z = np.average(self.vref_vs_tau[onwards:, 1])
        return Constants.convert(z, 'wavenumbers', to_AU=False) if ret_cm else z

    def window_avg(self, blocks=5, ret_cm=False):
        avgs = np.average(np.array_split(self.vref_vs_tau, blocks), axis=1)
        return Constants.convert(avgs, 'wavenumbers', to_AU=False) if ret_cm else avgs

    @staticmethod
    def get_training(training_file, ret_ang=False, ret_cm=False):
        """Walkers and energies of a training-data file (DEBUG_save_training_every)."""
        d = hdf5.read(training_file)
        cds, pots = d['coords'], d['pots']
        if ret_ang:
            cds = Constants.convert(cds, 'angstroms', to_AU=False)
        if ret_cm:
            pots = Constants.convert(pots, 'wavenumbers', to_AU=False)
        return cds, pots
