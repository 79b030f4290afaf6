"""MI355X-native (gfx950) quantized vector-sparse SpMM / SDDMM.

Device operators live in the in-tree library ``libmagicube_hip.so`` (HIP
kernels for gfx950 behind the C ABI in ``include/magicube_hip.h``); this
package is the host-side mirror of the reference's operator interfaces.
"""
from . import _lib
from ._lib import MC_SIGNED, MC_UNSIGNED, MagicubeError, lib

__all__ = ["MC_SIGNED", "MC_UNSIGNED", "MagicubeError", "lib"]
