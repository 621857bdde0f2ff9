"""PlanktonIndividuals.jl hot path on AMD Instinct MI355X.

Host-side mirror of the reference's user surface for the per-timestep path:
`RectilinearGrid`, `PlanktonModel(arch, grid; ...)`, `PlanktonSimulation`,
`update!` (here `update`), `TimeStep!` (here `TimeStep`), with the new
architecture `HIPArch` whose compute runs in libpi_hip.so (hand-written HIP for
gfx950).  Import this directory under any module name (its folder name is not a
Python identifier), e.g. with `importlib`.
"""
from ._abi import TRACER_NAMES
from .grids import RectilinearGrid, LatLonGrid, LoadLatLonGrid, Periodic, Bounded
from .params import (QuotaMode, CarbonMode, MacroMolecularMode, IronEnergyMode,
                     bgc_params_default, phyt_params_default, update_bgc_params, update_phyt_params,
                     abiotic_params_default, update_abiotic_params)
from .fields import default_tracer_init, generate_tracers, fill_halo_tracer
from .model import (Architecture, HIPArch, GPU, isfunctional, phyto_setup, PlanktonModel, ModelSpec,
                    TimeStep, initialize_plankton, set_bc, validate_bcs, BoundaryConditions, abiotic_setup, Palat,
                    set_bc_particle, initialize_abiotic_particle)
from .simulation import (PlanktonSimulation, PlanktonDiagnostics, update, default_PARF,
                         default_temperature)
from ._native import CapacityError, NativeError, load_library, lib_path
from .ensemble import Ensemble
from . import workloads
from . import autotune
from .output import PlanktonOutputWriter, write_output

update_ = update          # `update!`
TimeStep_ = TimeStep      # `TimeStep!`
write_output_ = write_output   # `write_output!`
set_bc_ = set_bc               # `set_bc!`
set_bc_particle_ = set_bc_particle   # `set_bc_particle!`

__all__ = ["RectilinearGrid", "LatLonGrid", "LoadLatLonGrid", "Periodic", "Bounded", "QuotaMode", "PlanktonModel", "HIPArch", "GPU",
           "phyto_setup", "PlanktonSimulation", "PlanktonDiagnostics", "update", "TimeStep",
           "bgc_params_default", "phyt_params_default", "update_bgc_params", "update_phyt_params",
           "default_tracer_init", "default_PARF", "default_temperature", "CapacityError",
           "TRACER_NAMES", "isfunctional", "Ensemble", "PlanktonOutputWriter", "write_output", "set_bc"]
