"""ctypes binding of ``libmagicube_hip.so`` (the C ABI in include/magicube_hip.h).

The library is built in-tree by ``python -m magicube_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for any operator.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmagicube_hip.so")

MC_OK = 0
MC_EINVAL_V = 1
MC_EINVAL_SHAPE = 2
MC_EINVAL_ARG = 3
MC_EHIP = 4
MC_ENOMEM = 5
MC_EUNSUPPORTED = 6

MC_SIGNED = 0
MC_UNSIGNED = 1

_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int

# name -> (restype, argtypes)
_SIGS = {
    "mc_version": (_I, []),
    "mc_status_string": (ctypes.c_char_p, [_I]),
    "mc_last_hip_error": (_I, []),
    "mc_spmm_num_tile_configs": (_I, []),
    "mc_spmm_get_tile_config": (_I, [_I, _P]),
    "mc_spmm_run": (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "mc_format_row_alignment": (_I, [_I, _I]),
    "mc_format_align_rows": (_I, [_I, _P, _I, _P]),
    "mc_format_align_cols": (_I, [_I, _P, _P, _P, _I, _I, _P]),
    "mc_format_prepare_values": (_I, [_I, _I, _I, _I, _P, _P, _I, _P, _P]),
    "mc_format_sorted_rows": (_I, [_I, _P, _P]),
}
_SIGS["mc_sddmm_l8r8"] = (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P])
_SIGS["mc_sddmm_l4r4"] = (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P])
_SIGS["mc_sddmm_l16r16"] = (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P])
_SIGS["mc_sddmm_run"] = (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P])
_F = ctypes.c_float
_LL = ctypes.c_longlong
_SIGS["mc_attn_quantize_batched"] = (_I, [_P, _I, _I, _I, _LL, _P])
_SIGS["mc_attn_absmax"] = (_I, [_P, _LL, _P, _P])
_SIGS["mc_attn_softmax_requant"] = (_I, [_I, _I, _P, _P, _P, _F, _F, _I, _P, _P, _P])
_SIGS["mc_attn_dequant"] = (_I, [_P, _I, _I, _F, _P, _P, _P])
_SIGS["mc_attn_forward"] = (_I, [_P, _P, _I, _I, _I, _I, _F, _I, _I, _I, _P])
_SIGS["mc_attn_dequant_f16"] = (_I, [_P, _LL, _F, _P, _P])
_SIGS["mc_layernorm_f16"] = (_I, [_LL, _I, _P, _P, _P, _P, _F, _P, _P])
_SIGS["mc_attn_softmax_f16"] = (_I, [_I, _I, _I, _P, _P, _P, _LL, _F, _F, _I, _P, _LL, _P])
_SIGS["mc_spmm_batched"] = (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _LL, _P, _LL, _P, _LL, _P])
_SIGS["mc_sddmm_batched"] = (_I, [_I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _LL, _P, _LL, _P, _LL, _I, _P])
for _name in ("mc_spmm_l8r8", "mc_spmm_l16r8", "mc_spmm_l8r4", "mc_spmm_l4r4", "mc_spmm_l12r8", "mc_spmm_l16r16",
              "mc_spmm_l12r4", "mc_spmm_l16r4"):
    _SIGS[_name] = (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P])


class MagicubeError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        msg = status_string(status) if _lib is not None else str(status)
        if status == MC_EHIP and _lib is not None:
            msg += f" (hipError {_lib.mc_last_hip_error()})"
        super().__init__(f"{where}: {msg} [status {status}]" if where else msg)
        self.status = status


def _register(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m magicube_amd.build` "
                "(the operators have no CPU fallback)")
        _lib = ctypes.CDLL(LIB_PATH)
        _register(_lib, _SIGS)
    return _lib


def register(sigs):
    """Declares further C signatures (used by the sddmm/attention modules)."""
    _SIGS.update(sigs)
    if _lib is not None:
        _register(_lib, sigs)


def status_string(status: int) -> str:
    return lib().mc_status_string(status).decode()


def check(status: int, where: str = "") -> None:
    if status != MC_OK:
        raise MagicubeError(status, where)


def exported_symbols():
    """Every mc_* name the library exports that include/*.h declares."""
    return sorted(_SIGS)
