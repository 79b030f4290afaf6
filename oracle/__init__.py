"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes access to ``oracle/liboracle.so`` (plain-C restatement of the
reference's host functions, see oracle.h for the citations) plus the FP64
attention oracle in :mod:`oracle.attention`.  Only tests/, the smoke check
and bench.py's cpu_baseline leg import this package, and only as the
checker; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double

_SIGS = {
    "orc_make_dense_matrix": (None, [_I, _L, _P, _I]),
    "orc_canonical_stream": (None, [_L, _P]),
    "orc_align_row_offsets": (_I, [_I, _P, _I, _P]),
    "orc_align_col_indices": (None, [_I, _P, _P, _P, _I, _P]),
    "orc_shuffle_col_indices": (None, [_I, _P, _P]),
    "orc_sorted_row_swizzle": (None, [_I, _P, _P]),
    "orc_prepare_values": (None, [_P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P]),
    "orc_spmm_ref_check": (_D, [_P, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I]),
    "orc_spmm_ref_check_omp": (_D, [_P, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I]),
    "orc_sddmm_ref_check": (_D, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I]),
    "orc_unpack_lanes": (None, [_P, _I, _L, _I, _P]),
    "orc_pack_lanes": (None, [_P, _I, _L, _P]),
    "orc_spmm_plain": (None, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "orc_sddmm_plain": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _I]),
}

_DT = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


LIB_V4_PATH = os.path.join(HERE, "liboracle_v4.so")
_AVX512 = ("avx512f", "avx512bw", "avx512dq", "avx512vl")


def _load(path):
    handle = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def host_flags() -> set:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def host_lib():
    """(library, march): the x86-64-v4 build where this CPU has AVX-512
    F/BW/DQ/VL, else the x86-64-v3 one.  Used for timing the host loops."""
    if os.path.exists(LIB_V4_PATH) and all(f in host_flags() for f in _AVX512):
        return _load(LIB_V4_PATH), "x86-64-v4"
    return lib(), "x86-64-v3"


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def make_dense_matrix(type_bytes: int, count: int, sddmm: bool = False) -> np.ndarray:
    out = np.empty(count, dtype=_DT[type_bytes])
    lib().orc_make_dense_matrix(type_bytes, count, _p(out), int(sddmm))
    return out


def canonical_stream(count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float32)
    lib().orc_canonical_stream(count, _p(out))
    return out


def align_row_offsets(row_offsets, align: int):
    ro = _i32(row_offsets)
    m_vec = ro.size - 1
    pairs = np.zeros(max(2 * m_vec, 1), dtype=np.int32)
    n = lib().orc_align_row_offsets(m_vec, _p(ro), align, _p(pairs))
    return pairs[: 2 * m_vec], int(n)


def align_col_indices(row_offsets, col_indices, pairs, aligned_num: int) -> np.ndarray:
    ro, ci, pr = _i32(row_offsets), _i32(col_indices), _i32(pairs)
    out = np.empty(max(aligned_num, 1), dtype=np.int32)
    lib().orc_align_col_indices(ro.size - 1, _p(ro), _p(ci), _p(pr), aligned_num, _p(out))
    return out[:aligned_num]


def shuffle_col_indices(aligned_cols) -> np.ndarray:
    ac = _i32(aligned_cols)
    out = np.empty(max(ac.size, 1), dtype=np.int32)
    lib().orc_shuffle_col_indices(ac.size, _p(ac), _p(out))
    return out[: ac.size]


def sorted_row_swizzle(row_offsets) -> np.ndarray:
    ro = _i32(row_offsets)
    out = np.empty(max(ro.size - 1, 1), dtype=np.int32)
    lib().orc_sorted_row_swizzle(ro.size - 1, _p(ro), _p(out))
    return out[: ro.size - 1]


def prepare_values(values: np.ndarray, type_bytes: int, scaleA: int, row_offsets, pairs, aligned_num: int,
                   V: int, preA: int, preA_cut: int, preB: int) -> np.ndarray:
    vals = np.ascontiguousarray(values)
    ro, pr = _i32(row_offsets), _i32(pairs)
    nbytes = aligned_num * scaleA * type_bytes
    out = np.zeros(max(nbytes, 1), dtype=np.uint8)
    lib().orc_prepare_values(_p(vals), type_bytes, scaleA, ro.size - 1, _p(ro), _p(pr), aligned_num, V, preA,
                             preA_cut, preB, _p(out))
    return out[:nbytes]


def spmm_ref_check(A: np.ndarray, type_bytes: int, B: np.ndarray, M: int, K: int, N: int, preA: int,
                   preA_cut: int, preB: int, V: int, row_offsets, col_indices, m_vec: int, scaleA: int,
                   handle=None):
    C = np.empty(M * N, dtype=np.int32)
    a = np.ascontiguousarray(A)
    b = _i32(B)
    ro, ci = _i32(row_offsets), _i32(col_indices)
    flops = (handle or lib()).orc_spmm_ref_check(_p(a), type_bytes, _p(b), _p(C), M, K, N, preA, preA_cut, preB, V, _p(ro),
                                     _p(ci), m_vec, scaleA)
    return C.reshape(M, N), flops


def spmm_ref_check_omp(A, type_bytes, B, M, K, N, preA, preB, V, row_offsets, col_indices, m_vec, threads=0,
                       handle=None):
    C = np.empty(M * N, dtype=np.int32)
    a = np.ascontiguousarray(A)
    b = _i32(B)
    ro, ci = _i32(row_offsets), _i32(col_indices)
    flops = (handle or lib()).orc_spmm_ref_check_omp(_p(a), type_bytes, _p(b), _p(C), M, K, N, preA, preB, V, _p(ro),
                                         _p(ci), m_vec, threads)
    return C.reshape(M, N), flops


def sddmm_ref_check(lhs, rhs, M, K, N, preA, preB, V, pairs, cols, m_vec, alignment, out_size):
    out = np.zeros(out_size, dtype=np.int32)
    l, r = _i32(lhs), _i32(rhs)
    pr, ci = _i32(pairs), _i32(cols)
    flops = lib().orc_sddmm_ref_check(_p(l), _p(r), _p(out), M, K, N, preA, preB, V, _p(pr), _p(ci), m_vec,
                                      alignment)
    return out, flops


def unpack_lanes(packed, bits: int, count: int, signed: bool) -> np.ndarray:
    p = np.ascontiguousarray(packed)
    out = np.empty(count, dtype=np.int32)
    lib().orc_unpack_lanes(_p(p), bits, count, int(signed), _p(out))
    return out


def pack_lanes(vals, bits: int) -> np.ndarray:
    v = _i32(vals)
    out = np.empty((v.size * bits + 7) // 8, dtype=np.uint8)
    lib().orc_pack_lanes(_p(v), bits, v.size, _p(out))
    return out


def spmm_plain(m_vec: int, V: int, N: int, row_offsets, col_indices, a, b) -> np.ndarray:
    ro, ci, av, bv = _i32(row_offsets), _i32(col_indices), _i32(a), _i32(b)
    C = np.empty(m_vec * V * N, dtype=np.int32)
    lib().orc_spmm_plain(m_vec, V, N, _p(ro), _p(ci), _p(av), _p(bv), _p(C))
    return C.reshape(m_vec * V, N)


def sddmm_plain(m_vec: int, V: int, K: int, pairs, cols, lhs, rhs, out_size: int, alignment: int) -> np.ndarray:
    pr, ci, l, r = _i32(pairs), _i32(cols), _i32(lhs), _i32(rhs)
    out = np.zeros(out_size, dtype=np.int32)
    lib().orc_sddmm_plain(m_vec, V, K, _p(pr), _p(ci), _p(l), _p(r), _p(out), alignment)
    return out
