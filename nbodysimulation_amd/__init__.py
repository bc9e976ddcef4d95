"""MI355X-native N-body engine: Barnes-Hut and brute-force gravity on gfx950.

The compute lives in libnbody.so (hand-written HIP kernels behind a C ABI,
include/nbody/nbody.h).  This package is the Python host mirror of the
reference's SimulationConfig/BarnesHutConfig API plus numpy bindings.
"""
from . import _lib
from ._lib import NbError, load
from .config import (BARNES_HUT, BRUTE_FORCE_GLOBAL, BRUTE_FORCE_LOCAL, EULER, FUNC_ZERO_VEL, LEAPFROG,
                     RANDOM_VEL, SPHERE_POS, STARTING_IN_VEL, STARTING_OUT_VEL, TANGENT_XZ_VEL, UNIFORM_POS,
                     BarnesHutConfig, EngineOptions, NormalDistribution, Simulation, SimulationConfig,
                     to_nb_config)
from .engine import NBodyEnsemble, NBodyServer, NBodySim, device_info, make_config, system_spec

__all__ = [
    "_lib", "NbError", "load", "NBodySim", "NBodyEnsemble", "NBodyServer", "system_spec", "make_config", "device_info", "SimulationConfig", "BarnesHutConfig",
    "Simulation", "NormalDistribution", "EngineOptions", "to_nb_config", "BARNES_HUT", "BRUTE_FORCE_GLOBAL",
    "BRUTE_FORCE_LOCAL", "EULER", "LEAPFROG", "SPHERE_POS", "UNIFORM_POS", "RANDOM_VEL", "STARTING_OUT_VEL",
    "STARTING_IN_VEL", "FUNC_ZERO_VEL", "TANGENT_XZ_VEL",
]
