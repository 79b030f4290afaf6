"""The reference's quantized sparse-transformer ops with their signatures.

Drop-in mirrors of the ``sptrans`` PyTorch extension the reference's layer
imports (end2end_eval/sparse_transformer_magicube/src/cuda/*.cpp): same
names, same positional arguments, same output dtypes and shapes, device
tensors in and out, every op running on the current stream (or ``stream``).

* ``quantization(x, bits, scale)`` / ``bquantization`` (quantization.cpp:3-20,
  quantization_kernel.cu:13-74,173-213): fp16 ``[..., L, D]`` -> int32
  ``[..., L, D * bits / 32]``, ``trunc(clamp(x * scale))``, lanes packed
  low first.
* ``sddmm_8b`` / ``sddmm_4b`` / ``bsddmm_8b`` / ``bsddmm_4b``
  (deq_sddmm.cpp:3-99, deq_sddmm_kernel.cu:447-466,565-636): packed Q
  ``[batch, m, k / (32 / bits)]`` and K ``[batch, n, ...]`` sampled at the
  shared mask -> fp16 ``[batch, nnz * V]`` (vector-major ``[slot][v]``),
  ``half(float(int32 S) / scale)``; padding slots hold 0.
* ``q_csr_softmax`` / ``q_bcsr_softmax`` (q_softmax.cpp:3-52,
  q_softmax_kernel.cu:180-271,394-479): fp16 logits -> int32 packed P of
  ``bits`` in the reference's SpMM A layouts (``mc_attn_softmax_f16``).
* ``bspmm_8b`` / ``bspmm_16b8b`` / ``bspmm_8b4b`` / ``bspmm_4b``
  (deq_spmm.cpp:5-141, deq_spmm_kernel.cu:593-876,1136-1361): packed P
  times packed V ``[batch, L, 64 * bits_rhs / 32]`` -> fp16
  ``[batch, m_vec * V, 64]``, ``half(float(int32 P V) / scale)``.

Masks are the reference's aligned vector-CSR: ``row_offsets`` 2 * m_vec
(start, end) pairs with starts on 16 slots (8-bit V) or 32 slots (4-bit V),
``column_indices`` -1 padded and not shuffled, ``row_indices`` the row order
(any permutation).  The products are exact two's-complement integers: the
reference feeds its signed quantized Q/K/V to unsigned tensor-core MMAs
(quantization_kernel.cu:65-69 against sddmm_utils/compute_utils.h:102),
which changes its results wherever a lane is negative; these ops compute the
signed products the layer means.

4-bit V runs on the library's 4-bit-B SpMM kernels, whose A operand is the
natural nibble packing with column indices shuffled within groups of 8
([0, 2, 4, 6, 1, 3, 5, 7], SpMM/SpMM/spmm_benchmark.cpp:204-222): the ops
shuffle the mask's indices, and re-lay 8-bit P into nibble planes and 16-bit
P into byte planes, with device tensor ops before the launch.
"""
from __future__ import annotations

import contextlib

import torch

from . import _lib
from .attention import bquantization as _bquantization

_SHUFFLE8 = (0, 2, 4, 6, 1, 3, 5, 7)


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_handle(stream):
    import ctypes
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _on(stream):
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def _i32(t, what):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{what} must be a device tensor")
    if t.dtype != torch.int32:
        raise ValueError(f"{what} must be int32, not {t.dtype}")
    return t.contiguous()


def _check_v(vec_length):
    if vec_length not in (2, 4, 8):
        raise ValueError(f"vec_length must be 2, 4 or 8, not {vec_length}")


def _mask(row_indices, row_offsets, column_indices):
    ro = _i32(row_offsets, "row_offsets")
    ci = _i32(column_indices, "column_indices")
    ri = _i32(row_indices, "row_indices") if row_indices is not None else None
    if ro.numel() % 2:
        raise ValueError("row_offsets holds (start, end) pairs")
    m_vec = ro.numel() // 2
    if ri is not None and ri.numel() != m_vec:
        raise ValueError(f"row_indices has {ri.numel()} entries for {m_vec} vector rows")
    return ri, ro, ci, m_vec


# ---------------------------------------------------------------------------
# quantization.cpp
# ---------------------------------------------------------------------------

def bquantization(input_matrix: torch.Tensor, bits: int, scale: float, stream=None) -> torch.Tensor:
    """fp16 [batch, L, D] -> int32 [batch, L, D * bits / 32], trunc(clamp(x * scale))."""
    return _bquantization(input_matrix, bits, scale, stream=stream)


def quantization(input_matrix: torch.Tensor, bits: int, scale: float, stream=None) -> torch.Tensor:
    """fp16 [L, D] -> int32 [L, D * bits / 32] (quantization.cpp:3-8)."""
    return _bquantization(input_matrix, bits, scale, stream=stream)


# ---------------------------------------------------------------------------
# deq_sddmm.cpp
# ---------------------------------------------------------------------------

def _bsddmm(bits, row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length, scale, stream,
            batched):
    _check_v(vec_length)
    ri, ro, ci, m_vec = _mask(row_indices, row_offsets, column_indices)
    lhs = _i32(lhs_matrix, "lhs_matrix")
    rhs = _i32(rhs_matrix, "rhs_matrix")
    if batched:
        if lhs.dim() != 3 or rhs.dim() != 3 or lhs.shape[0] != rhs.shape[0]:
            raise ValueError("lhs_matrix [batch, m, k/(32/bits)] and rhs_matrix [batch, n, ...] expected")
        batch = rhs.shape[0]
    else:
        if lhs.dim() != 2 or rhs.dim() != 2:
            raise ValueError("lhs_matrix [m, k/(32/bits)] and rhs_matrix [n, ...] expected")
        batch = 1
    m, k_int32 = lhs.shape[-2], lhs.shape[-1]
    n = rhs.shape[-2]
    k = k_int32 * (32 // bits)
    if rhs.shape[-1] != k_int32:
        raise ValueError("lhs_matrix and rhs_matrix must share k")
    if m != m_vec * vec_length:
        raise ValueError(f"lhs_matrix has {m} rows; the mask covers {m_vec} vector rows of {vec_length}")
    nnz = ci.numel()
    with _on(stream):
        acc = torch.zeros((batch, nnz * vec_length), dtype=torch.int32, device=lhs.device)
        out = torch.empty((batch, nnz * vec_length) if batched else (nnz * vec_length,), dtype=torch.float16,
                          device=lhs.device)
        lib, st = _lib.lib(), _stream_handle(stream)
        _lib.check(lib.mc_sddmm_batched(bits, _lib.MC_SIGNED, 1, batch, m_vec, k, n, _ptr(ri), _ptr(ro), _ptr(ci),
                                        _ptr(lhs), m * k_int32 * 4, _ptr(rhs), n * k_int32 * 4, _ptr(acc),
                                        nnz * vec_length, vec_length, st), f"bsddmm_{bits}b")
        _lib.check(lib.mc_attn_dequant_f16(_ptr(acc), acc.numel(), float(scale), _ptr(out), st), "bsddmm dequant")
    return out


def bsddmm_8b(row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length: int, bits: int,
              scale: float, stream=None) -> torch.Tensor:
    """Batched S = Q K^T at the mask, fp16(int32 / scale) (deq_sddmm.cpp:82-93)."""
    return _bsddmm(8, row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length, scale, stream,
                   True)


def bsddmm_4b(row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length: int, bits: int,
              scale: float, stream=None) -> torch.Tensor:
    """4-bit Q and K (deq_sddmm.cpp:56-67)."""
    return _bsddmm(4, row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length, scale, stream,
                   True)


def sddmm_8b(row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length: int, bits: int,
             scale: float, stream=None) -> torch.Tensor:
    """One matrix: lhs [m, k/4], rhs [n, k/4] -> fp16 [nnz * V] (deq_sddmm.cpp:69-80)."""
    return _bsddmm(8, row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length, scale, stream,
                   False)


def sddmm_4b(row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length: int, bits: int,
             scale: float, stream=None) -> torch.Tensor:
    """One matrix, 4-bit operands (deq_sddmm.cpp:43-54)."""
    return _bsddmm(4, row_indices, row_offsets, column_indices, lhs_matrix, rhs_matrix, vec_length, scale, stream,
                   False)


# ---------------------------------------------------------------------------
# q_softmax.cpp
# ---------------------------------------------------------------------------

def q_bcsr_softmax(row_indices, row_offsets, values, sqrt_dk: float, scale: float, vec_length: int, batch_size: int,
                   bits: int, stream=None) -> torch.Tensor:
    """fp16 logits [batch, nnz * V] -> int32 [batch, nnz * V * bits / 32]: P
    quantized to ``bits`` (16, 8 or 4) in the reference's layouts, zeros
    between a row's padded end and the next row (q_softmax.cpp:37-49)."""
    _check_v(vec_length)
    if bits not in (16, 8, 4):
        raise ValueError("bits must be 16, 8 or 4")
    ri = _i32(row_indices, "row_indices")
    ro = _i32(row_offsets, "row_offsets")
    if not isinstance(values, torch.Tensor) or values.dtype != torch.float16 or not values.is_cuda:
        raise ValueError("values must be a float16 device tensor")
    vals = values.contiguous()
    if batch_size < 1 or vals.numel() % batch_size:
        raise ValueError(f"{vals.numel()} values do not split into {batch_size} matrices")
    stride = vals.numel() // batch_size
    if (stride * bits) % 32:
        raise ValueError(f"{stride} values per matrix do not pack into int32 words of {bits}-bit P")
    attn_stride = stride * bits // 32
    with _on(stream):
        attn = torch.zeros((batch_size, attn_stride), dtype=torch.int32, device=vals.device)
        _lib.check(_lib.lib().mc_attn_softmax_f16(batch_size, ri.numel(), vec_length, _ptr(ri), _ptr(ro), _ptr(vals),
                                                  stride, float(sqrt_dk), float(scale), bits, _ptr(attn),
                                                  attn_stride * 4, _stream_handle(stream)), "q_bcsr_softmax")
    return attn


def q_csr_softmax(row_indices, row_offsets, values, sqrt_dk: float, scale: float, vec_length: int, bits: int,
                  stream=None) -> torch.Tensor:
    """One matrix: fp16 [nnz * V] -> int32 [nnz * V * bits / 32] (q_softmax.cpp:12-22)."""
    return q_bcsr_softmax(row_indices, row_offsets, values, sqrt_dk, scale, vec_length, 1, bits, stream)[0]


# ---------------------------------------------------------------------------
# deq_spmm.cpp
# ---------------------------------------------------------------------------

def _shuffled(column_indices: torch.Tensor) -> torch.Tensor:
    """Column indices shuffled within groups of 8, as the 4-bit-B kernels read them."""
    if column_indices.numel() % 8:
        raise ValueError("4-bit V needs the mask aligned to 32 slots")
    perm = torch.tensor(_SHUFFLE8, dtype=torch.long, device=column_indices.device)
    return column_indices.view(-1, 8)[:, perm].reshape(-1).contiguous()


def _p8_to_nibble_planes(p: torch.Tensor, batch: int, V: int) -> torch.Tensor:
    """8-bit P [(i/16)][v][16] bytes -> per 32-slot group [low nibbles][V][16
    bytes] + [high nibbles][V][16 bytes] (the L8-R4 operand)."""
    b = p.view(torch.uint8).view(batch, -1, 2, V, 16).permute(0, 1, 3, 2, 4).reshape(batch, -1, V, 32)
    lo, hi = b & 15, b >> 4
    pack = lambda x: x[..., 0::2] | (x[..., 1::2] << 4)          # noqa: E731
    return torch.stack([pack(lo), pack(hi)], dim=2).contiguous().view(torch.int32).view(batch, -1)


def _p16_to_byte_planes(p: torch.Tensor, batch: int, V: int) -> torch.Tensor:
    """16-bit P [(i/16)][v][low 16][high 16] -> [(i/16)][low, high][V][16]
    (the L16-R8 operand)."""
    b = p.view(torch.uint8).view(batch, -1, V, 2, 16).permute(0, 1, 3, 2, 4)
    return b.contiguous().view(torch.int32).view(batch, -1)


def _bspmm(bits_lhs, bits_rhs, row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length, scale,
           stream):
    _check_v(vec_length)
    ri, ro, ci, m_vec = _mask(row_indices, row_offsets, column_indices)
    vals = _i32(values, "values")
    rhs = _i32(rhs_matrix, "rhs_matrix")
    if rhs.dim() != 3:
        raise ValueError("rhs_matrix [batch, k, n * bits / 32] expected")
    batch, k, n_int32 = rhs.shape
    n = n_int32 * (32 // bits_rhs)
    if vals.numel() % batch:
        raise ValueError("values do not split into one P per matrix")
    vals = vals.view(batch, -1)
    nnz = ci.numel()
    need = nnz * vec_length * bits_lhs // 8
    if vals.shape[1] * 4 < need:
        raise ValueError(f"values hold {vals.shape[1] * 4} bytes per matrix; the mask needs {need}")
    with _on(stream):
        if bits_rhs == 4:
            ci = _shuffled(ci)
        if bits_lhs == 8 and bits_rhs == 4:
            vals = _p8_to_nibble_planes(vals[:, : need // 4], batch, vec_length)
        elif bits_lhs == 16:
            vals = _p16_to_byte_planes(vals[:, : need // 4], batch, vec_length)
        m = m_vec * vec_length
        acc = torch.empty((batch, m, n), dtype=torch.int32, device=rhs.device)
        out = torch.empty((batch, m, n), dtype=torch.float16, device=rhs.device)
        lib, st = _lib.lib(), _stream_handle(stream)
        _lib.check(lib.mc_spmm_batched(bits_lhs, bits_rhs, _lib.MC_SIGNED, -1, batch, m_vec, vec_length, n, k,
                                       _ptr(ri), _ptr(ro), _ptr(ci), _ptr(vals), vals.stride(0) * 4, _ptr(rhs),
                                       k * n_int32 * 4, _ptr(acc), m * n, st), f"bspmm L{bits_lhs}-R{bits_rhs}")
        _lib.check(lib.mc_attn_dequant_f16(_ptr(acc), acc.numel(), float(scale), _ptr(out), st), "bspmm dequant")
    return out


def bspmm_8b(row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length: int, bits_lhs: int,
             bits_rhs: int, scale: float, stream=None) -> torch.Tensor:
    """8-bit P x 8-bit V -> fp16 [batch, m, 64] (deq_spmm.cpp:73-94)."""
    return _bspmm(8, 8, row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length, scale, stream)


def bspmm_16b8b(row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length: int, bits_lhs: int,
                bits_rhs: int, scale: float, stream=None) -> torch.Tensor:
    """16-bit P x 8-bit V (deq_spmm.cpp:119-141)."""
    return _bspmm(16, 8, row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length, scale, stream)


def bspmm_8b4b(row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length: int, bits_lhs: int,
               bits_rhs: int, scale: float, stream=None) -> torch.Tensor:
    """8-bit P x 4-bit V (deq_spmm.cpp:96-117)."""
    return _bspmm(8, 4, row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length, scale, stream)


def bspmm_4b(row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length: int, bits_lhs: int,
             bits_rhs: int, scale: float, stream=None) -> torch.Tensor:
    """4-bit P x 4-bit V (deq_spmm.cpp:50-71)."""
    return _bspmm(4, 4, row_indices, row_offsets, column_indices, values, rhs_matrix, vec_length, scale, stream)
