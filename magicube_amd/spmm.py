"""SpMM operators on device tensors (host mirror of spmm::wmmaSpmm_*).

``wmmaSpmm_8b`` / ``_12b8b`` / ``_16b8b`` / ``_16b`` / ``_8b4b`` / ``_12b4b`` /
``_16b4b`` / ``_4b`` take the reference's arguments in the reference's order
(SpMM/SpMM/include/wmma_spmm.cuh:15-78) as torch tensors resident in HBM,
plus the lane interpretation and a stream, and raise ``MagicubeError`` where
the reference returns a non-success ``cudaError_t``.  12-bit lanes sit in
16-bit containers.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .formats import PreparedSpmm

SIGNED = _lib.MC_SIGNED
UNSIGNED = _lib.MC_UNSIGNED

_PAIR_FN = {(8, 8): "mc_spmm_l8r8", (16, 8): "mc_spmm_l16r8", (8, 4): "mc_spmm_l8r4", (4, 4): "mc_spmm_l4r4",
            (12, 8): "mc_spmm_l12r8", (16, 16): "mc_spmm_l16r16", (12, 4): "mc_spmm_l12r4",
            (16, 4): "mc_spmm_l16r4"}
PAIRS = tuple(_PAIR_FN)


def _dptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    if not t.is_cuda:
        raise ValueError("operands must be device tensors")
    if not t.is_contiguous():
        raise ValueError("operands must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _check_bytes(t, need: int, what: str) -> None:
    """A tensor operand must hold at least ``need`` bytes (raw integer
    pointers are the caller's responsibility, as in the C ABI)."""
    if isinstance(t, torch.Tensor) and t.numel() * t.element_size() < need:
        raise ValueError(f"{what} holds {t.numel() * t.element_size()} bytes; the launch reads or writes {need}")


def _check_device(tensors, what: str) -> None:
    """Every tensor operand of one launch lives on one device."""
    devs = {t.device for t in tensors if isinstance(t, torch.Tensor)}
    if any(d.type != "cuda" for d in devs):
        raise ValueError(f"{what}: operands must be device tensors")
    if len(devs) > 1:
        raise ValueError(f"{what}: operands are on several devices {sorted(map(str, devs))}")


def wmma_spmm(lhs_bits: int, rhs_bits: int, m_vec: int, vec_length: int, n: int, k: int, row_indices,
              row_offsets, column_indices, values, rhs_matrix, output_matrix, sign: int = SIGNED,
              config: int | None = None, stream=None) -> None:
    if min(m_vec, vec_length, n, k) >= 0:
        _check_bytes(output_matrix, m_vec * vec_length * n * 4, "output_matrix")
        _check_bytes(rhs_matrix, k * n * rhs_bits // 8, "rhs_matrix")
    _check_device((row_indices, row_offsets, column_indices, values, rhs_matrix, output_matrix), "spmm")
    lib = _lib.lib()
    args = (m_vec, vec_length, n, k, _dptr(row_indices), _dptr(row_offsets), _dptr(column_indices),
            _dptr(values), _dptr(rhs_matrix), _dptr(output_matrix))
    if config is None:
        fn = getattr(lib, _PAIR_FN[(lhs_bits, rhs_bits)])
        st = fn(*args, sign, _stream(stream))
    else:
        st = lib.mc_spmm_run(lhs_bits, rhs_bits, sign, config, *args, _stream(stream))
    _lib.check(st, f"spmm L{lhs_bits}-R{rhs_bits}")


def wmmaSpmm_8b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(8, 8, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_16b8b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                   output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(16, 8, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_8b4b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                  output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(8, 4, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_4b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(4, 4, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_12b8b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                   output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(12, 8, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_16b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                 output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(16, 16, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_12b4b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                   output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(12, 4, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def wmmaSpmm_16b4b(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
                   output_matrix, sign=SIGNED, stream=None):
    wmma_spmm(16, 4, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values, rhs_matrix,
              output_matrix, sign, None, stream)


def tile_configs():
    lib = _lib.lib()
    out = []
    for i in range(lib.mc_spmm_num_tile_configs()):
        buf = (ctypes.c_int * 6)()
        _lib.check(lib.mc_spmm_get_tile_config(i, buf), "mc_spmm_get_tile_config")
        out.append(tuple(buf))
    return out


@dataclass
class DeviceSpmm:
    """A prepared sparse operand resident in HBM, ready for repeated products."""

    prep: PreparedSpmm
    row_offsets: torch.Tensor
    col_indices: torch.Tensor
    values: torch.Tensor
    row_indices: torch.Tensor
    # tile config for products with a given n (the autotuner's choice;
    # absent: the library's default for the shape)
    configs: dict = field(default_factory=dict)

    @classmethod
    def upload(cls, prep: PreparedSpmm, device="cuda") -> "DeviceSpmm":
        def t(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device)

        # never hand the library a NULL pointer for an empty array
        vals = prep.values
        pad = max(16 - vals.size, (-vals.size) % 16)
        if pad:
            vals = np.concatenate([vals, np.zeros(pad, np.uint8)])
        cols = prep.col_indices
        if cols.size == 0:
            cols = np.full(16, -1, np.int32)
        ro = prep.row_offsets if prep.row_offsets.size else np.zeros(2, np.int32)
        ri = prep.row_indices if prep.row_indices.size else np.zeros(1, np.int32)
        return cls(prep, t(ro), t(cols), t(vals), t(ri))

    def matmul(self, rhs: torch.Tensor, n: int, out: torch.Tensor | None = None, sign: int = SIGNED,
               config: int | None = None, stream=None) -> torch.Tensor:
        p = self.prep
        if out is None:
            out = torch.empty((p.m_vec * p.V, n), dtype=torch.int32, device=rhs.device)
        if config is None:
            config = self.configs.get(n)
        wmma_spmm(p.lhs_bits, p.rhs_bits, p.m_vec, p.V, n, p.k, self.row_indices, self.row_offsets,
                  self.col_indices, self.values, rhs, out, sign, config, stream)
        return out
