"""pyvibdmc_amd: the pyvibdmc DMC propagation loop on one MI355X (gfx950).

Drop-in names of the reference package (``import pyvibdmc as pv`` scripts keep working with
``import pyvibdmc_amd as pv``): DMC_Sim, dmc_restart, Potential, Potential_NoMP,
NN_Potential, Potential_Direct, ImpSampManager, ImpSampManager_NoMP, Constants.
Results: SimInfo (the reference's reader of the HDF5 outputs).
Device potentials and trial wave functions (fused into the step kernels):
HarmonicOscillator1D, MorseOscillator1D, MorseBendTriatomic, TwoMonomerCluster,
HarmonicTrial1D, GaussianProductTrial.  Surrogate potential: SurrogatePotential.
"""
from .constants import Constants, get_atomic_num, get_atomic_string
from .potentials import (Potential, Potential_NoMP, NN_Potential, Potential_Direct, DevicePotential,
                         HarmonicOscillator1D, MorseOscillator1D, MorseBendTriatomic, TwoMonomerCluster,
                         PartridgeSchwenkeWater)
from .impsamp import (ImpSamp, ImpSampManager, ImpSampManager_NoMP, DeviceTrial, HarmonicTrial1D,
                      GaussianProductTrial, water_trial, water_trial_widths, cluster6_trial)
from .sim import DMC_Sim, dmc_restart
from .surrogate import SurrogateModel, SurrogatePotential, like_atoms
from ._native import DMCError, load as load_library
from .analysis import SimInfo
from .initial_conditions import HarmonicAnalysis, InitialConditioner, MolFiniteDifference
from .hdf5 import save_h5
from . import torch_ops

__version__ = '0.1.0'
