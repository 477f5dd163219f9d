"""rimu_amd -- an MI355X (gfx950) FCIQMC step engine behind Rimu.jl's interfaces.

The package mirrors the part of Rimu.jl's API that drives the FCIQMC step:
``ProjectorMonteCarloProblem`` / ``init`` / ``solve`` / ``lomc`` (``lomc!``), the
``AbstractDVec`` and ``StochasticStyle`` contracts, and the Hamiltonian interface of
``HubbardReal1D``, ``HubbardMom1D`` and ``HubbardRealSpace`` (BoseFS, FermiFS and
two-component CompositeFS addresses on CubicGrid geometries).  All compute goes through ``librimu_hip.so``
(``include/rimu_hip.h``); the package has no CPU fallback.

Julia names with ``!`` map to: ``solve!`` -> ``PMCSimulation.solve_``, ``step!`` ->
``PMCSimulation.step_``, ``lomc!`` -> ``lomc``.
"""
from . import _native, statstools
from ._native import RimuError, CapacityError, annihilate, device_count
from .addresses import BoseFS, CompositeFS, FermiFS, FermiFS2C, near_uniform
from .geometry import CubicGrid, HardwallBoundaries, LadderBoundaries, PeriodicBoundaries
from .hamiltonians import (HubbardMom1D, HubbardReal1D, HubbardRealSpace, diagonal_element,
                           dimension, get_offdiagonal, num_offdiagonals, offdiagonals,
                           starting_address, hubbard_dispersion)
from .styles import (IsDeterministic, IsDynamicSemistochastic, IsStochasticInteger,
                     NoCompression, ThresholdCompression, WithReplacement, default_style)
from .initiators import (CoherentInitiator, Initiator, InitiatorRule, InitiatorValue,
                         NonInitiator, SimpleInitiator)
from .dictvectors import DVec, DeviceDVec, FrozenDVec, InitiatorDVec, PDVec, freeze
from .strategies import (ConstantTimeStep, DoubleLogUpdate, DoubleLogUpdateAfterTargetWalkers,
                         DontUpdate, LogUpdate, LogUpdateAfterTargetWalkers, ProjectedEnergy)
from .pmc import (PMCSimulation, ProjectorMonteCarloProblem, default_starting_vector, init,
                  lomc, solve)
from .observables import DensityMatrixDiagonal, G2RealCorrelator, dot
from .replicas import AllOverlaps, NoStats
from .dense import DenseHamiltonians, basis_size
from . import rimuio as RimuIO
from .rimuio import load_df, load_state, save_df, save_state
from .batched import BatchedRuns, RunSpec, u_sweep_specs

__all__ = [n for n in dir() if not n.startswith("_")]
