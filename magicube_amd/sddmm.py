"""SDDMM operators on device tensors (host mirror of sddmm::wmmaSddmm_*).

``wmmaSddmm_8b`` / ``wmmaSddmm_4b`` / ``wmmaSddmm_16b`` take the reference's arguments in the
reference's order (SDDMM/SDDMM/include/wmma_sddmm.cuh:7-32) as torch tensors
resident in HBM, plus the lane interpretation, the output layout and a stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, formats
from .spmm import _check_bytes, _check_device, _dptr, _stream

SIGNED = _lib.MC_SIGNED
UNSIGNED = _lib.MC_UNSIGNED
OUT_REFERENCE = 0   # [(j/8)][v][j%8]
OUT_VECTOR = 1      # [j][v]


def wmma_sddmm(bits: int, m_vec: int, k: int, n: int, row_indices, row_offsets, col_indices, lhs_matrix,
               rhs_matrix, output_values, vec_length: int, sign: int = SIGNED, layout: int = OUT_REFERENCE,
               stream=None) -> None:
    if min(m_vec, k, n, vec_length) >= 0:
        _check_bytes(lhs_matrix, m_vec * vec_length * k * bits // 8, "lhs_matrix")
        _check_bytes(rhs_matrix, n * k * bits // 8, "rhs_matrix")
        # every slot of [start, roundup8(end)) is written; the column-index
        # array (-1 padded) holds at least as many slots as the rows reach
        if isinstance(col_indices, torch.Tensor):
            _check_bytes(output_values, -(-col_indices.numel() // 8) * 8 * vec_length * 4, "output_values")
    _check_device((row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values), "sddmm")
    st = _lib.lib().mc_sddmm_run(bits, sign, layout, m_vec, k, n, _dptr(row_indices), _dptr(row_offsets),
                                 _dptr(col_indices), _dptr(lhs_matrix), _dptr(rhs_matrix), _dptr(output_values),
                                 vec_length, _stream(stream))
    _lib.check(st, f"sddmm {bits}-bit")


def wmmaSddmm_8b(m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
                 vec_length, sign=SIGNED, stream=None):
    wmma_sddmm(8, m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
               vec_length, sign, OUT_REFERENCE, stream)


def wmmaSddmm_4b(m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
                 vec_length, sign=SIGNED, stream=None):
    wmma_sddmm(4, m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
               vec_length, sign, OUT_REFERENCE, stream)


def wmmaSddmm_16b(m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
                  vec_length, sign=SIGNED, stream=None):
    wmma_sddmm(16, m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix, rhs_matrix, output_values,
               vec_length, sign, OUT_REFERENCE, stream)


@dataclass
class DeviceMask:
    """A vector-sparse sampling mask resident in HBM (pairs aligned to ``alignment``)."""

    m_vec: int
    n: int
    V: int
    aligned_num: int
    pairs: np.ndarray
    cols: np.ndarray
    row_offsets: torch.Tensor
    col_indices: torch.Tensor
    row_indices: torch.Tensor

    @classmethod
    def upload(cls, csr: formats.VectorCSR, alignment: int = 8, device="cuda") -> "DeviceMask":
        pairs, an = formats.align_rows(csr.row_offsets, alignment)
        cols = formats.align_cols(csr.row_offsets, csr.col_indices, pairs, an, shuffle=False)
        order = formats.sorted_rows(csr.row_offsets)

        def t(a, fill):
            a = np.ascontiguousarray(a) if a.size else np.full(16, fill, np.int32)
            return torch.from_numpy(a).to(device)

        return cls(csr.m_vec, csr.k, csr.V, an, pairs, cols, t(pairs, 0), t(cols, -1), t(order, 0))

    def output_size(self) -> int:
        """int32 values the launch may write: the padded slot array rounded up to 8, times V."""
        return -(-max(self.aligned_num, self.col_indices.numel()) // 8) * 8 * self.V

    def sddmm(self, bits: int, lhs: torch.Tensor, rhs: torch.Tensor, k: int, out: torch.Tensor | None = None,
              sign: int = SIGNED, layout: int = OUT_REFERENCE, stream=None) -> torch.Tensor:
        if out is None:
            out = torch.zeros(self.output_size(), dtype=torch.int32, device=lhs.device)
        wmma_sddmm(bits, self.m_vec, k, self.n, self.row_indices, self.row_offsets, self.col_indices, lhs, rhs,
                   out, self.V, sign, layout, stream)
        return out
