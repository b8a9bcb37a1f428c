"""palabos_amd: MI355X-native D3Q19 collide-and-stream engine behind Palabos's lattice API.

The compute path is the native library palabos_amd/libpalabos_amd.so (HIP for
gfx950, C ABI in include/palabos_amd.h). Importing this package does not load
it; the first lattice does, and fails loudly if it is missing.
"""
from .lattice import (D3Q19, BGK, RegularizedBGK, BounceBack, NoDynamics, MomentumExchangeBounceBack,
                      WallLink, Bulk, Block3D)
from .native import NativeLibraryError, PlbError

__all__ = ["D3Q19", "BGK", "RegularizedBGK", "BounceBack", "NoDynamics", "MomentumExchangeBounceBack",
           "WallLink", "Bulk", "Block3D", "NativeLibraryError", "PlbError"]
