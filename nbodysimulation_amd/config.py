"""Python mirror of the reference's configuration types.

Same class names, getter/setter names, enum values and defaults as
src/SimulationConfig.h:10-133 and src/BarnesHutConfig.h:10-61, so code written
against the reference's API reads the same here.  The C++ drop-in of the same
types is include/nbody/config.hpp.  ``to_nb_config`` lowers a SimulationConfig
to the C ABI's nb_config, adding the engine's own fields (rule set, law,
accumulation, dt).
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

from . import _lib as L

# enums (src/SimulationConfig.h:10-38)
LEAPFROG, EULER = L.LEAPFROG, L.EULER
SPHERE_POS, UNIFORM_POS = L.SPHERE_POS, L.UNIFORM_POS
RANDOM_VEL, STARTING_OUT_VEL, STARTING_IN_VEL, FUNC_ZERO_VEL, TANGENT_XZ_VEL = (
    L.RANDOM_VEL, L.STARTING_OUT_VEL, L.STARTING_IN_VEL, L.FUNC_ZERO_VEL, L.TANGENT_XZ_VEL)
BARNES_HUT, BRUTE_FORCE_GLOBAL, BRUTE_FORCE_LOCAL = L.BARNES_HUT, L.BRUTE_FORCE_GLOBAL, L.BRUTE_FORCE_LOCAL


@dataclass
class NormalDistribution:
    """src/SimulationConfig.h:40-44."""
    mean: float
    deviation: float


class BarnesHutConfig:
    """src/BarnesHutConfig.h:10-61: max octree depth 32, theta 1.5, epsilon 0.01."""

    def __init__(self, other: "BarnesHutConfig | None" = None):
        self.max_octree_depth = 32
        self.theta = 1.5
        self.epsilon = 0.01
        if other is not None:
            self.max_octree_depth, self.theta, self.epsilon = other.max_octree_depth, other.theta, other.epsilon

    def GetMaxOctreeDepth(self) -> int: return self.max_octree_depth
    def GetTheta(self) -> float: return self.theta
    def GetEpsilon(self) -> float: return self.epsilon
    def SetMaxOctreeDepth(self, v: int) -> None: self.max_octree_depth = int(v)
    def SetTheta(self, v: float) -> None: self.theta = float(v)
    def SetEpsilon(self, v: float) -> None: self.epsilon = float(v)


class SimulationConfig:
    """src/SimulationConfig.h:46-105 with its defaults (N=20000, G=6.67e-11,
    speed 1.0, radius 1.5, no massive objects, UNIFORM_POS, FUNC_ZERO_VEL,
    BRUTE_FORCE_GLOBAL, mass Normal(0.5, 0.25), EULER, no collision)."""

    def __init__(self, other: "SimulationConfig | None" = None):
        if other is not None:
            self.__dict__.update(copy.deepcopy(other.__dict__))
            return
        self.num_of_bodies = 20000
        self.gravitational_constant = 6.67e-11
        self.starting_speed_mul = 1.0
        self.starting_volume_radius = 1.5
        self.number_of_massive_objects = 0
        self.massive_object_mass = 1.0
        self.collision = False
        self.pos_config = UNIFORM_POS
        self.vel_config = FUNC_ZERO_VEL
        self.algo_config = BRUTE_FORCE_GLOBAL
        self.mass_distr = NormalDistribution(0.5, 0.25)
        self.barnes_hut_config = BarnesHutConfig()
        self.numerical_method = EULER

    def GetNumberOfBodies(self) -> int: return self.num_of_bodies
    def GetGravitationalConstant(self) -> float: return self.gravitational_constant
    def GetStartingSpeedMul(self) -> float: return self.starting_speed_mul
    def GetStartingVolumeRadius(self) -> float: return self.starting_volume_radius
    def GetNumberOfMassiveObjects(self) -> int: return self.number_of_massive_objects
    def GetMassiveObjectMass(self) -> float: return self.massive_object_mass
    def GetPositionConfig(self) -> int: return self.pos_config
    def GetVelocityConfig(self) -> int: return self.vel_config
    def GetMassDistribution(self) -> NormalDistribution: return self.mass_distr
    def GetAlgorithmConfig(self) -> int: return self.algo_config
    def GetBarnesHutConfig(self) -> BarnesHutConfig: return self.barnes_hut_config
    def GetNumericalMethod(self) -> int: return self.numerical_method
    def GetCollision(self) -> bool: return self.collision

    def SetNumberOfBodies(self, v: int) -> None: self.num_of_bodies = int(v)
    def SetGravitationalConstant(self, v: float) -> None: self.gravitational_constant = float(v)
    def SetStartingSpeedMul(self, v: float) -> None: self.starting_speed_mul = float(v)
    def SetStartingVolumeRadius(self, v: float) -> None: self.starting_volume_radius = float(v)
    def SetNumberOfMassiveObjects(self, v: int) -> None: self.number_of_massive_objects = int(v)
    def SetMassiveObjectMass(self, v: float) -> None: self.massive_object_mass = float(v)
    def SetPositionConfig(self, v: int) -> None: self.pos_config = int(v)
    def SetVelocityConfig(self, v: int) -> None: self.vel_config = int(v)
    def SetMassDistribution(self, v: NormalDistribution) -> None: self.mass_distr = NormalDistribution(v.mean, v.deviation)
    def SetAlgorithmConfig(self, v: int) -> None: self.algo_config = int(v)
    def SetBarnesHutConfig(self, v: BarnesHutConfig) -> None: self.barnes_hut_config = BarnesHutConfig(v)
    def SetNumericalMethod(self, v: int) -> None: self.numerical_method = int(v)
    def SetCollision(self, v: bool) -> None: self.collision = bool(v)


class Simulation:
    """src/SimulationConfig.h:107-133: GetConfig returns a copy."""

    def __init__(self):
        self._config = SimulationConfig()
        self.simulationSpeedMul = 1.0
        self.particleSize = 0.05

    def GetConfig(self) -> SimulationConfig: return SimulationConfig(self._config)
    def SetConfig(self, config: SimulationConfig) -> None: self._config = SimulationConfig(config)
    def GetSimulationSpeedMul(self) -> float: return self.simulationSpeedMul
    def GetParticleSize(self) -> float: return self.particleSize


@dataclass
class EngineOptions:
    """Engine choices the reference does not expose (documented in nbody.h)."""
    rules: int = L.RULES_REFERENCE
    law: int = L.LAW_REFERENCE_UNIT
    accumulation: int = L.ACCUM_PLAIN
    dt: float = 1e-3
    cell_capacity: int = 0
    use_graph: bool = True
    time_stages: bool = False


def to_nb_config(cfg: SimulationConfig, opts: EngineOptions | None = None) -> L.NbConfig:
    opts = opts or EngineOptions()
    lib = L.load()
    c = L.NbConfig()
    lib.nb_config_default(c)
    c.num_bodies = cfg.GetNumberOfBodies()
    c.algorithm = cfg.GetAlgorithmConfig()
    c.method = cfg.GetNumericalMethod()
    c.rules = opts.rules
    c.law = opts.law
    c.accumulation = opts.accumulation
    c.max_octree_depth = cfg.GetBarnesHutConfig().GetMaxOctreeDepth()
    c.cell_capacity = opts.cell_capacity
    c.gravitational_constant = cfg.GetGravitationalConstant()
    c.theta = cfg.GetBarnesHutConfig().GetTheta()
    c.epsilon = cfg.GetBarnesHutConfig().GetEpsilon()
    c.dt = opts.dt
    c.collision = 1 if cfg.GetCollision() else 0
    c.use_graph = 1 if opts.use_graph else 0
    c.time_stages = 1 if opts.time_stages else 0
    return c
