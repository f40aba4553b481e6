"""nebula_amd: MI355X-native electron-transport hot path of the Nebula SEM
simulator (octree traversal, Kieft-Bosch / boundary physics, particle
management) as HIP kernels behind a C ABI (include/nbl_hip.h).

The HIP library must be built in-tree (nebula_amd/_build/libnbl_hip.so);
importing the package fails if it is missing.
"""
from . import _lib
from ._lib import NblError
from .api import Geometry, Materials, PreparedMaterial, Sampler, octree_linearize, prepared, device_info

__all__ = ["NblError", "Geometry", "Materials", "PreparedMaterial", "Sampler", "octree_linearize", "prepared",
           "device_info"]
