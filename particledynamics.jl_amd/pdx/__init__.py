"""pdx — MI355X-native force hot path of ParticleDynamics.jl (host side over libpdx's C ABI).

Importing this package loads ``particledynamics.jl_amd/lib/libpdx.so``; it fails loudly
when the library is missing. Device arrays are torch tensors on ``cuda`` (HIP).
"""
from ._lib import (ArgumentError, DeviceError, DimensionMismatch, NeighborCapacityError, check, last_error,
                   lib, LIB_PATH)
from .neighbors import DeviceNeighborWorkspace, CellList, build_neighbors_dense
from .forces import (LennardJones, WCA, SoftRepulsive, MixedSigmaLJ, MixedSigmaWCA, PairMatrixLJ, PairMatrixWCA,
                     RealSpaceCoulomb, PairForce, FENEBonds, DomainError)
from .electrostatics import (ElectrostaticsSpec, PME, PPPM, ESP, esp_rule, ElectrostaticsSolver, alpha_from_tolerance,
                             fft_friendly, ewald_fp64)
from .engine import ForceEngine, default_capacity
from .replicas import ReplicaBatch
from .clusters import ClusterPairEngine
from . import systems

__all__ = ["ArgumentError", "DeviceError", "DimensionMismatch", "NeighborCapacityError", "check", "last_error",
           "lib", "LIB_PATH", "DeviceNeighborWorkspace", "CellList", "build_neighbors_dense", "LennardJones", "WCA",
           "SoftRepulsive", "MixedSigmaLJ", "MixedSigmaWCA", "PairMatrixLJ", "PairMatrixWCA", "RealSpaceCoulomb", "PairForce", "FENEBonds", "DomainError", "ElectrostaticsSpec", "PME", "PPPM", "ESP", "esp_rule",
           "ElectrostaticsSolver",
           "alpha_from_tolerance", "fft_friendly", "ewald_fp64", "ForceEngine", "default_capacity", "ReplicaBatch", "ClusterPairEngine", "systems"]
