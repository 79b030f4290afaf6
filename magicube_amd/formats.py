"""Host-side vector-CSR preparation (mirror of BmFN's preparation steps).

A *plain* vector-sparse matrix is a CSR over V x 1 column vectors:
``row_offsets`` (m_vec + 1), ``col_indices`` (nnz_vec) and the packed
vectors (nnz_vec x V lanes of ``lhs_bits``).  :func:`prepare_spmm` turns it
into the layouts the device operators take: (start, end) row pairs padded to
the pair's row alignment, padded/shuffled column indices, prepared values
and the longest-row-first order (SpMM/SpMM/spmm_benchmark.cpp:177-394).
The work is done by the library's C++ ``mc_format_*`` functions.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class VectorCSR:
    """Plain vector-sparse pattern: m_vec vector rows of V rows each, k columns."""

    m_vec: int
    k: int
    V: int
    row_offsets: np.ndarray  # int32, m_vec + 1
    col_indices: np.ndarray  # int32, nnz_vec

    @property
    def nnz_vec(self) -> int:
        return int(self.row_offsets[-1])

    @property
    def m(self) -> int:
        return self.m_vec * self.V


@dataclass
class PreparedSpmm:
    """Device-ready (host-resident) SpMM operand in the reference layouts."""

    lhs_bits: int
    rhs_bits: int
    V: int
    m_vec: int
    k: int
    nnz_vec: int
    aligned_num: int
    row_offsets: np.ndarray   # int32, 2*m_vec (start, end)
    col_indices: np.ndarray   # int32, aligned_num (-1 padding; shuffled for 4-bit B)
    values: np.ndarray        # uint8, aligned_num * V * container_bits(lhs_bits) / 8
    row_indices: np.ndarray   # int32, m_vec


def container_bits(bits: int) -> int:
    """Storage width of a lane: 12-bit lanes sit in 16-bit containers."""
    return 16 if bits == 12 else bits


def row_alignment(lhs_bits: int, rhs_bits: int) -> int:
    return int(_lib.lib().mc_format_row_alignment(lhs_bits, rhs_bits))


def align_rows(row_offsets: np.ndarray, alignment: int):
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    m_vec = ro.size - 1
    pairs = np.zeros(max(2 * m_vec, 1), dtype=np.int32)
    n = _lib.lib().mc_format_align_rows(m_vec, _ptr(ro), alignment, _ptr(pairs))
    if n < 0:
        raise ValueError("invalid row offsets")
    return pairs[: 2 * m_vec], int(n)


def align_cols(row_offsets, col_indices, pairs, aligned_num: int, shuffle: bool) -> np.ndarray:
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(col_indices, dtype=np.int32)
    pr = np.ascontiguousarray(pairs, dtype=np.int32)
    out = np.empty(max(aligned_num, 1), dtype=np.int32)
    _lib.check(_lib.lib().mc_format_align_cols(ro.size - 1, _ptr(ro), _ptr(ci), _ptr(pr), aligned_num,
                                               int(shuffle), _ptr(out)), "mc_format_align_cols")
    return out[:aligned_num]


def prepare_values(lhs_bits: int, rhs_bits: int, V: int, row_offsets, pairs, aligned_num: int,
                   values: np.ndarray) -> np.ndarray:
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    pr = np.ascontiguousarray(pairs, dtype=np.int32)
    vals = np.ascontiguousarray(values).view(np.uint8)
    nbytes = aligned_num * V * container_bits(lhs_bits) // 8
    out = np.empty(max(nbytes, 1), dtype=np.uint8)
    _lib.check(_lib.lib().mc_format_prepare_values(lhs_bits, rhs_bits, V, ro.size - 1, _ptr(ro), _ptr(pr),
                                                   aligned_num, _ptr(vals), _ptr(out)),
               "mc_format_prepare_values")
    return out[:nbytes]


def sorted_rows(row_offsets) -> np.ndarray:
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    out = np.empty(max(ro.size - 1, 1), dtype=np.int32)
    _lib.check(_lib.lib().mc_format_sorted_rows(ro.size - 1, _ptr(ro), _ptr(out)), "mc_format_sorted_rows")
    return out[: ro.size - 1]


def prepare_spmm(csr: VectorCSR, values: np.ndarray, lhs_bits: int, rhs_bits: int,
                 sort_rows: bool = True) -> PreparedSpmm:
    """Pads, shuffles and lays out a plain vector-CSR for the SpMM operators."""
    align = row_alignment(lhs_bits, rhs_bits)
    pairs, aligned_num = align_rows(csr.row_offsets, align)
    cols = align_cols(csr.row_offsets, csr.col_indices, pairs, aligned_num, shuffle=(rhs_bits == 4))
    vals = prepare_values(lhs_bits, rhs_bits, csr.V, csr.row_offsets, pairs, aligned_num, values)
    order = sorted_rows(csr.row_offsets) if sort_rows else np.arange(csr.m_vec, dtype=np.int32)
    return PreparedSpmm(lhs_bits, rhs_bits, csr.V, csr.m_vec, csr.k, csr.nnz_vec, aligned_num, pairs, cols,
                        vals, order)


# ---- packed lanes (little-endian, lane 0 in the low bits) -----------------

def pack_lanes(vals: np.ndarray, bits: int) -> np.ndarray:
    """int lanes -> packed bytes (values truncated to ``bits``; 12-bit lanes
    in 16-bit containers)."""
    v = np.ascontiguousarray(vals).astype(np.int64).ravel()
    if bits == 12:
        return (v & 0xFFF).astype(np.uint16).view(np.uint8)
    if bits == 8:
        return (v & 0xFF).astype(np.uint8)
    if bits == 16:
        return (v & 0xFFFF).astype(np.uint16).view(np.uint8)
    if bits == 4:
        if v.size % 2:
            v = np.concatenate([v, np.zeros(1, np.int64)])
        lo = v[0::2] & 0xF
        hi = v[1::2] & 0xF
        return (lo | (hi << 4)).astype(np.uint8)
    raise ValueError(f"unsupported lane width {bits}")


def unpack_lanes(packed: np.ndarray, bits: int, count: int, signed: bool) -> np.ndarray:
    """packed bytes -> int32 lanes."""
    b = np.ascontiguousarray(packed).view(np.uint8).ravel()
    if bits == 12:
        out = (b[: 2 * count].view(np.uint16) & 0xFFF).astype(np.int32)
    elif bits == 8:
        out = b[:count].astype(np.int32)
    elif bits == 16:
        out = b[: 2 * count].view(np.uint16).astype(np.int32)
    elif bits == 4:
        out = np.empty(b.size * 2, dtype=np.int32)
        out[0::2] = b & 0xF
        out[1::2] = b >> 4
        out = out[:count]
    else:
        raise ValueError(f"unsupported lane width {bits}")
    if signed:
        top = 1 << (bits - 1)
        out = ((out ^ top) - top).astype(np.int32)
    return out
