"""MI355X-native AVBD rigid-body step: a drop-in for the reference's WebGPU GpuSolver3D.

The step runs as hand-written HIP kernels for gfx950 in libavbd_mi355x.so behind a C ABI
(include/avbd.h); ``GpuSolver3D`` here is the host-side mirror of the reference's class over it.
"""
from .solver import (GpuSolver3D, StorageCapacityError, default_params, lib, FLAG_DEFAULTS, FLAG_MATCH_NEAREST,
                     FLAG_FACE_BIAS, FLAG_REUSE_CONTACTS, FLAG_START_AT_REST, FLAG_MASS_PENALTY, NO_COLOR, STICK_BIT,
                     EXPORTS, morton_order, estimate_pairs, plan_capacity)
from .sdf import AnalyticField, VoxelField, SdfCollisionBridge, collision_params

__all__ = ["GpuSolver3D", "StorageCapacityError", "default_params", "lib", "EXPORTS", "morton_order", "estimate_pairs",
           "plan_capacity", "AnalyticField", "VoxelField", "SdfCollisionBridge", "collision_params"]
