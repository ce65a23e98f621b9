"""fornos_amd -- MI355X-native baking hot path of Fornos.

The compute runs in libfornos_hip.so (hand-written gfx950 HIP kernels behind
the C ABI of include/fornos_hip.h); this package is its Python binding and
the seeded bake workloads.
"""
from . import _abi
from ._abi import FornosError, SYMBOLS, lib
from .api import (AmbientOcclusionSolver, Batch, BentNormalsSolver, CompressedMapUV, HeightSolver, Mesh,
                  MeshMapping, NormalsSolver, PositionSolver, Scene, Solver, TexelArrays, ThicknessSolver,
                  device_count, sample_table)
from .configs import CONFIGS, prepare

__all__ = ["AmbientOcclusionSolver", "Batch", "BentNormalsSolver", "CompressedMapUV", "HeightSolver", "Mesh",
           "MeshMapping", "NormalsSolver", "PositionSolver", "Scene", "Solver", "TexelArrays", "ThicknessSolver",
           "device_count", "sample_table", "CONFIGS", "prepare", "FornosError", "SYMBOLS", "lib"]
