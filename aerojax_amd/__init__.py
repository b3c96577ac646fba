"""aerojax_amd -- MI355X (gfx950) implementation of AeroJAX's collocated
projection-method Navier-Stokes timestep.

Public API (reference signatures):
  rk_step_unified        advection_schemes/rk3_simple_new.py:175-180
  poisson_multigrid      pressure_solvers/multigrid_solver.py:55-56
  step_pure              solver/pure_functions/step.py:49-66
  step_collocated        solver/simulation/step_handlers.py:62-185 (GUI step)
  get_re_parameters      solver/params.py:308-413
plus ``poisson_multigrid_to_tolerance`` (textbook multigrid, FP64 defect
correction) and ``CollocatedStepper`` (preallocated repeated stepping).
All arithmetic runs in hand-written HIP kernels in ``libaerojax_hip.so``.
"""
from ._lib import AeroJaxError, AeroJaxLibraryError, load as load_library, version
from .advection_schemes import rk_step_unified
from .operators import (constant_smagorinsky, divergence, divergence_nonperiodic,
                        dynamic_smagorinsky, pressure_rhs, project, project_collocated,
                        smagorinsky_total)
from .batch import BatchedStepper
from .diagnostics import flow_diagnostics
from .params import get_re_parameters
from .pressure_solvers import (MultigridPlan, poisson_multigrid,
                               poisson_multigrid_to_tolerance)
from .snapshot import (DatasetWriter, Snapshot, load_dataset, pairs_from_dataset,
                       take_snapshot)
from .step import CollocatedStepper, SimState, initial_state, step_collocated, step_pure

__all__ = [
    "AeroJaxError", "AeroJaxLibraryError", "load_library", "version", "rk_step_unified",
    "constant_smagorinsky", "dynamic_smagorinsky", "divergence", "divergence_nonperiodic", "pressure_rhs", "project",
    "project_collocated", "smagorinsky_total", "get_re_parameters", "flow_diagnostics",
    "MultigridPlan", "poisson_multigrid", "poisson_multigrid_to_tolerance",
    "BatchedStepper", "CollocatedStepper", "SimState", "initial_state", "step_collocated", "step_pure",
    "Snapshot", "take_snapshot", "DatasetWriter", "load_dataset", "pairs_from_dataset",
]
