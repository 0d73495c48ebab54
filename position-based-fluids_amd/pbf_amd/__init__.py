"""MI355X-native PBF substep: Python host binding of libpbf_amd.so."""
from ._lib import (PBF_FLAG_FRIENDS_OVERFLOW, PBF_FLAG_NAN, PBF_FLAG_SORT_TIMEOUT, LIB_PATH, LibraryMissing,
                   PbfError, PbfOptions, PbfParams, lib)
from .scenes import (Scene, config, dam_break, dam_break_block, default_options, morton_grid_size, physical_params,
                     reference_block)
from .simulation import BatchSimulation, Simulation, load_parameters
from .batch import ensemble

__all__ = [
    "BatchSimulation", "ensemble", "LIB_PATH", "LibraryMissing", "PbfError", "PbfOptions", "PbfParams", "Scene", "Simulation", "config",
    "dam_break", "dam_break_block", "default_options", "lib", "load_parameters", "morton_grid_size", "reference_block",
    "physical_params", "PBF_FLAG_FRIENDS_OVERFLOW", "PBF_FLAG_NAN", "PBF_FLAG_SORT_TIMEOUT",
]
