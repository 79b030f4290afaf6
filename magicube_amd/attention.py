"""Quantized sparse attention (host mirror of the reference's spMultiheadAttention
core, end2end_eval/sparse_transformer_magicube/spattention.py:182-216).

An :class:`AttentionBatch` holds any number of (sample, head) matrices of
uneven sequence length, each with its own vector-sparse mask, resident in
HBM: fp16 inputs, quantized operands, masks, outputs and the device
descriptors of the fused launch.  It runs

* :meth:`AttentionBatch.forward` — one batched quantize launch per operand
  width (K and V; Q too when the fused kernel does not read it as fp16) plus
  ONE persistent fused launch over a longest-first work list of (matrix,
  vector row) items (capturable in a hipGraph), or
* :meth:`AttentionBatch.forward_plain` — the one-op-per-launch pipeline per
  matrix (quantize x3, SDDMM, softmax + requantize, SpMM, dequantize).

Both give identical int32 logits, int8 probabilities and int32 P V.
:func:`bquantization` and the int32-logit :func:`softmax_requant` sit at the
end of the module; the reference's op signatures (``bsddmm_8b``,
``q_bcsr_softmax``, ``bspmm_8b`` ...) are in :mod:`magicube_amd.sptrans`.
"""
from __future__ import annotations

import contextlib
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, formats
from .spmm import _dptr, _stream

HEAD_DIM = 64
# the kernels address Q/K/V rows with 32-bit byte offsets and mark masked
# slots with offsets >= 2^31
MAX_SEQ_LEN = 1 << 24
P_REFERENCE = 0
P_ROW_SCALED = 1
P_REFERENCE_4BIT = 2        # P_REFERENCE with P clamped at 7 (4-bit P)
ROUND_TRUNC = 0
ROUND_NEAREST = 1
QUANT_RANGE4 = 2            # or-ed into a rounding mode: 4-bit values stored as int8
MAX_QUANT_JOBS = 65535      # jobs of one mc_attn_quantize_batched launch
SIGNED = _lib.MC_SIGNED


class _AttnMatrixC(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("m_vec", ctypes.c_int), ("row_offsets", ctypes.c_void_p),
                ("col_indices", ctypes.c_void_p), ("q", ctypes.c_void_p), ("k", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("alpha", ctypes.c_float), ("o_scale", ctypes.c_float),
                ("o", ctypes.c_void_p), ("o_int", ctypes.c_void_p), ("p_out", ctypes.c_void_p),
                ("s_out", ctypes.c_void_p), ("o16", ctypes.c_void_p), ("o16_stride", ctypes.c_int),
                ("o16_pad", ctypes.c_int), ("q16", ctypes.c_void_p), ("q16_stride", ctypes.c_int),
                ("q_scale", ctypes.c_float), ("q_rounding", ctypes.c_int), ("q_pad", ctypes.c_int)]


class _QuantJobC(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("count", ctypes.c_longlong), ("scale", ctypes.c_float),
                ("out", ctypes.c_void_p), ("row_len", ctypes.c_longlong), ("row_stride", ctypes.c_longlong)]


def _rows(x: torch.Tensor) -> tuple[int, int]:
    """(row_len, row_stride) of a quantize job's source: (0, 0) when x is
    contiguous, else the rows of a [L, 64] view with unit column stride."""
    if x.is_contiguous():
        return 0, 0
    if x.dim() != 2 or x.stride(1) != 1 or x.shape[1] % 8 != 0:
        raise ValueError("q, k, v must be contiguous or [L, 64] views with unit column stride")
    return int(x.shape[1]), int(x.stride(0))


def _as_input(x: torch.Tensor) -> torch.Tensor:
    """A q / k / v operand as the quantize launch reads it: kept as the
    caller's tensor when contiguous or a row-strided [L, 64] view, else
    copied."""
    try:
        _rows(x)
        return x
    except ValueError:
        return x.contiguous()


def _bytes_tensor(raw: bytes, device) -> torch.Tensor:
    """Host bytes -> device uint8 tensor, copied asynchronously from pinned
    memory (the caching host allocator keeps the staging buffer alive until
    the copy has run)."""
    device = torch.device(device)
    host = torch.frombuffer(bytearray(raw if raw else bytes(16)), dtype=torch.uint8)
    if device.type == "cuda":
        host = host.pin_memory()
    return host.to(device, non_blocking=True)


def _pinned(a: np.ndarray) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.pin_memory() if torch.cuda.is_available() else t


def _struct_tensor(structs, device) -> torch.Tensor:
    return _bytes_tensor(b"".join(bytes(s) for s in structs), device)


@dataclass
class ResidentMask:
    """A vector-sparse attention mask in the kernels' device format: (start,
    end) row pairs with every row padded to a multiple of 16 slots, column
    indices padded with -1, a longest-row-first launch order and the per-row
    nonzero counts."""

    L: int
    V: int
    m_vec: int
    nnz_vec: int
    aligned_num: int
    row_offsets: torch.Tensor   # int32 [2 * m_vec]
    col_indices: torch.Tensor   # int32 [max(aligned_num, 16)]
    row_indices: torch.Tensor   # int32 [max(m_vec, 1)]
    row_nnz: torch.Tensor       # int32 [max(m_vec, 1)]
    pairs: np.ndarray | None = None   # host copy of row_offsets when built from a host mask

    @classmethod
    def from_csr(cls, mask: formats.VectorCSR, device) -> "ResidentMask":
        pairs, an = formats.align_rows(mask.row_offsets, 16)
        cols = formats.align_cols(mask.row_offsets, mask.col_indices, pairs, an, shuffle=False)
        order = formats.sorted_rows(mask.row_offsets)
        nnz_rows = np.diff(mask.row_offsets).astype(np.int32)
        dev = lambda a, fill: torch.from_numpy(a if a.size else fill).to(device)  # noqa: E731
        return cls(L=mask.k, V=mask.V, m_vec=mask.m_vec, nnz_vec=mask.nnz_vec, aligned_num=an,
                   row_offsets=dev(pairs, np.zeros(2, np.int32)),
                   col_indices=dev(cols, np.full(16, -1, np.int32)),
                   row_indices=dev(order, np.zeros(1, np.int32)),
                   row_nnz=dev(nnz_rows, np.zeros(1, np.int32)), pairs=pairs)

    def hbm_bytes(self) -> int:
        return sum(t.numel() * 4 for t in (self.row_offsets, self.col_indices, self.row_indices, self.row_nnz))


@dataclass
class AttentionMatrix:
    """One (sample, head) problem: fp16 q, k, v of shape [L, 64] and a V-row
    mask (host :class:`formats.VectorCSR` or device :class:`ResidentMask`)."""

    q: torch.Tensor
    k: torch.Tensor
    v: torch.Tensor
    mask: formats.VectorCSR | ResidentMask

    @property
    def L(self) -> int:
        return int(self.q.shape[0])


@dataclass
class _Dev:
    L: int
    m_vec: int
    nnz: int
    aligned_num: int
    pairs: np.ndarray | None
    row_offsets: torch.Tensor
    col_indices: torch.Tensor
    row_indices: torch.Tensor
    row_nnz: torch.Tensor
    q16: torch.Tensor
    k16: torch.Tensor
    v16: torch.Tensor
    qq: torch.Tensor
    kq: torch.Tensor
    vq: torch.Tensor
    o: torch.Tensor
    scales: tuple
    alpha: float
    o_scale: float
    qk_bits: int
    v_bits: int = 8
    o_int: torch.Tensor | None = None
    p_out: torch.Tensor | None = None
    s_out: torch.Tensor | None = None
    # plain-path work buffers (allocated on first use)
    work: dict = field(default_factory=dict)

    def q16_row_stride(self) -> int | None:
        """Row stride (in halves) at which the fused kernel can read Q as
        fp16 and quantize it itself, or None: rows 16-byte aligned and the
        whole extent addressable by the kernel's 32-bit offsets."""
        x = self.q16
        if x.dtype != torch.float16 or not x.is_cuda or x.data_ptr() % 16:
            return None
        if x.is_contiguous():
            stride = HEAD_DIM
        elif x.dim() == 2 and x.shape[1] == HEAD_DIM and x.stride(1) == 1:
            stride = int(x.stride(0))
        else:
            return None
        if stride % 8 or (self.L - 1) * stride + HEAD_DIM >= 1 << 30:
            return None
        return stride

    def desc_bytes(self, q_rounding: int | None = None) -> bytes:
        """The launch descriptor; with ``q_rounding`` (the Q/K quantize
        rounding mode) the kernel reads Q as fp16 and quantizes it itself."""
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        half = self.o.dtype == torch.float16
        fq = q_rounding is not None
        return bytes(_AttnMatrixC(self.L, self.m_vec, self.row_offsets.data_ptr(), self.col_indices.data_ptr(),
                                  self.qq.data_ptr(), self.kq.data_ptr(), self.vq.data_ptr(), self.alpha,
                                  self.o_scale, 0 if half else self.o.data_ptr(), ptr(self.o_int),
                                  ptr(self.p_out), ptr(self.s_out), self.o.data_ptr() if half else 0,
                                  int(self.o.stride(0)) if half else 0, 0,
                                  self.q16.data_ptr() if fq else 0, self.q16_row_stride() if fq else 0,
                                  float(self.scales[0]) if fq else 0.0, int(q_rounding) if fq else 0, 0))

    def quant_job_bytes(self) -> tuple[bytes, bytes]:
        """(q and k jobs, v job) of the batched quantize launches."""
        qk = b"".join(bytes(_QuantJobC(x.data_ptr(), x.numel(), s, y.data_ptr(), *_rows(x)))
                      for x, s, y in ((self.q16, self.scales[0], self.qq), (self.k16, self.scales[1], self.kq)))
        return qk, bytes(_QuantJobC(self.v16.data_ptr(), self.v16.numel(), self.scales[2], self.vq.data_ptr(),
                                    *_rows(self.v16)))


def working_set_bytes(L: int, qk_bits: int, v_bits: int = 8) -> int:
    """Device bytes a matrix allocates when it joins a batch: quantized q, k,
    v and the fp32 output."""
    qk_row = HEAD_DIM * qk_bits // 8
    return 2 * max(L * qk_row, 16) + max(L * HEAD_DIM * v_bits // 8, 16) + L * HEAD_DIM * 4


def make_resident(m: AttentionMatrix, scales: tuple, qk_bits: int, V: int, p_scale: float, device,
                  debug: bool = False, out: torch.Tensor | None = None, v_bits: int = 8) -> _Dev:
    """Uploads and aligns the mask when it is a host mask, and allocates the
    matrix's quantized operands and output, or writes into ``out``: a
    contiguous fp32 [L, 64] tensor, or an fp16 [L, 64] view with unit column
    stride and a row stride that is a multiple of 4 (8-byte aligned).  q, k
    and v may be row-strided [L, 64] views (read in place)."""
    mask = m.mask if isinstance(m.mask, ResidentMask) else ResidentMask.from_csr(m.mask, device)
    if m.q.shape[1] != HEAD_DIM or mask.V != V or mask.m_vec * V != m.L:
        raise ValueError("matrix shapes do not match the batch")
    if m.L > MAX_SEQ_LEN:
        raise ValueError(f"sequence length {m.L} exceeds {MAX_SEQ_LEN}")
    qk_row = HEAD_DIM * qk_bits // 8
    if out is not None:
        ok32 = out.dtype == torch.float32 and out.is_contiguous()
        ok16 = (out.dtype == torch.float16 and out.stride(1) == 1 and out.stride(0) % 4 == 0
                and out.data_ptr() % 8 == 0 and out.stride(0) < 2 ** 31)
        if tuple(out.shape) != (m.L, HEAD_DIM) or not (ok32 or ok16):
            raise ValueError("out must be a contiguous float32 [L, 64] tensor or an fp16 [L, 64] view with unit "
                             "column stride, a row stride that is a multiple of 4 and 8-byte alignment")
    sq, sk, sv = (float(x) for x in scales)
    alpha = float(np.float32(1.0) / (np.float32(sq) * np.float32(sk) * np.float32(math.sqrt(HEAD_DIM))))
    o_scale = float(np.float32(1.0) / (np.float32(sv) * np.float32(p_scale)))
    d = _Dev(L=m.L, m_vec=mask.m_vec, nnz=mask.nnz_vec, aligned_num=mask.aligned_num, pairs=mask.pairs,
             row_offsets=mask.row_offsets, col_indices=mask.col_indices, row_indices=mask.row_indices,
             row_nnz=mask.row_nnz, q16=_as_input(m.q), k16=_as_input(m.k), v16=_as_input(m.v),
             qq=torch.empty(max(m.L * qk_row, 16), dtype=torch.uint8, device=device),
             kq=torch.empty(max(m.L * qk_row, 16), dtype=torch.uint8, device=device),
             vq=torch.empty(max(m.L * HEAD_DIM * v_bits // 8, 16), dtype=torch.uint8, device=device),
             o=out if out is not None else torch.empty((m.L, HEAD_DIM), dtype=torch.float32, device=device),
             scales=(sq, sk, sv), alpha=alpha, o_scale=o_scale, qk_bits=qk_bits, v_bits=v_bits)
    if debug:
        an = mask.aligned_num
        d.o_int = torch.empty((m.L, HEAD_DIM), dtype=torch.int32, device=device)
        d.p_out = torch.zeros(max(an * V, 16), dtype=torch.uint8, device=device)
        d.s_out = torch.zeros(max(an * V, 16), dtype=torch.int32, device=device)
    return d


def tensor_scales(tensors, stream=None) -> list[float]:
    """127 / max|x| per fp16 tensor, max computed on the GPU in one pass each."""
    lib = _lib.lib()
    dev = tensors[0].device
    bits = torch.zeros(len(tensors), dtype=torch.int32, device=dev)
    for i, t in enumerate(tensors):
        _lib.check(lib.mc_attn_absmax(_dptr(t), t.numel(), ctypes.c_void_p(bits.data_ptr() + 4 * i),
                                      _stream(stream)), "mc_attn_absmax")
    amax = bits.cpu().view(torch.float32).numpy()
    return [127.0 / float(a) if a > 0 else 1.0 for a in amax]


# 4-bit Q/K clip at this many RMS: a uniform 16-level quantizer of N(0, 1)
# data has its smallest squared error near a 2.5 sigma clip, and the layer's
# error against FP64 attention is 0.16 there, against 0.25 at 7 / max|x|
# (L = 2048, 80 % sparse, V = 8)
QK4_CLIP_RMS = 2.5


def qkv_scales(mats, qk_bits: int = 8) -> list[tuple]:
    """Static per-tensor quantization scales of every matrix's (Q, K, V):
    127 / max|x| for 8-bit operands (and V), 7 / (2.5 rms(x)) for 4-bit Q
    and K."""
    flat = tensor_scales([t for m in mats for t in (m.q, m.k, m.v)])
    if qk_bits == 4:
        with torch.no_grad():
            qk = [t for m in mats for t in (m.q, m.k)]
            rms = torch.stack([t.float().pow(2).mean() for t in qk]).sqrt().cpu().numpy()
        for n, r in enumerate(rms):
            flat[3 * (n // 2) + n % 2] = 7.0 / (QK4_CLIP_RMS * float(r)) if r > 0 else 1.0
    return [tuple(flat[3 * i:3 * i + 3]) for i in range(len(mats))]


N_XCD = 8              # MI355X: workgroups are handed to the 8 XCDs in turn
WAVES_PER_BLOCK = 4    # waves per workgroup of the fused attention launch
# 0: one wave per work item.  With four waves per SIMD resident, letting the
# dispatcher hand out items beat a persistent grid of 8192 waves by 3.5-6 %
# (32 matrices of L = 4096; 128 matrices of L ~ U[512, 4096]).
DEFAULT_GRID_WAVES = 0


def xcd_bins(m_vecs, works, n_xcd: int = N_XCD) -> list[np.ndarray]:
    """Rows of the matrices dealt to each XCD: n_xcd int32 [k, 2] arrays of
    (matrix, vector row).  Matrices sorted by ``works`` (largest first) go to
    the XCDs in boustrophedon order (0..7, 7..0, ...), which balances the
    XCDs' work without a per-matrix loop."""
    m_vecs = np.asarray(m_vecs, dtype=np.int64)
    works = np.asarray(works, dtype=np.float64)
    es = np.argsort(-works, kind="stable")
    es = es[m_vecs[es] > 0]
    k = np.arange(es.size)
    lane = k % n_xcd
    bins = np.where((k // n_xcd) % 2 == 0, lane, n_xcd - 1 - lane)
    lists = []
    for x in range(n_xcd):
        ex = es[bins == x]
        mv = m_vecs[ex]
        e_col = np.repeat(ex, mv).astype(np.int32)
        r_col = (np.arange(int(mv.sum())) - np.repeat(np.cumsum(mv) - mv, mv)).astype(np.int32)
        lists.append(np.stack([e_col, r_col], axis=1))
    return lists


def interleave_bins(lists, wpb: int = WAVES_PER_BLOCK) -> np.ndarray:
    """Work list of the fused launch from per-XCD row lists: the rows of XCD
    x fill the workgroups b = x, x + n, x + 2n, ... (wpb items each), which
    the hardware hands to XCD x, so each matrix's K/V rows are read through
    one XCD's L2.  Slots left over in the shorter lists hold (-1, -1) and are
    skipped by the kernel."""
    n_xcd = len(lists)
    blocks = max((-(-len(li) // wpb) for li in lists), default=0)
    if blocks == 0:
        return np.full((1, 2), -1, dtype=np.int32)
    padded = np.full((n_xcd, blocks * wpb, 2), -1, dtype=np.int32)
    for x, li in enumerate(lists):
        padded[x, :len(li)] = li
    # [xcd][block][wave] -> [block][xcd][wave]: workgroup b * n_xcd + x on XCD x
    return np.ascontiguousarray(padded.reshape(n_xcd, blocks, wpb, 2).transpose(1, 0, 2, 3)).reshape(-1, 2)


def xcd_item_order(m_vecs, works, n_xcd: int = N_XCD, wpb: int = WAVES_PER_BLOCK) -> np.ndarray:
    """Work list of the fused launch: int32 [n, 2] (matrix, vector row), each
    matrix's rows on one XCD (:func:`xcd_bins`, :func:`interleave_bins`)."""
    return interleave_bins(xcd_bins(m_vecs, works, n_xcd), wpb)


class AttentionBatch:
    """Device-resident quantized sparse attention over uneven matrices."""

    def __init__(self, mats: list[AttentionMatrix], qk_bits: int = 8, V: int = 8, p_mode: int = P_ROW_SCALED,
                 p_scale: float = 127.0, scales: list[tuple] | None = None, debug: bool = False,
                 rounding: int | None = None, device="cuda", resident: list[_Dev] | None = None,
                 outputs: list[torch.Tensor] | None = None, v_range4: bool = False, v_bits: int = 8,
                 fuse_q: bool = True):
        """v_bits 4: V is quantized to 4 bits and held packed ([L][32] bytes),
        as the reference's rhs_pre 4 layer holds it (Q/K must be 4-bit too).
        v_range4: 8-bit V storage with values clamped to the 4-bit range
        [-8, 7] (the same integers, one per byte).
        fuse_q: the fused kernel reads Q as fp16 and quantizes it itself
        (the same integers as the quantize launch; the quantize launch then
        covers K and V only, and the quantized Q buffer is left unwritten by
        the fused path).  Taken when every matrix's Q is a CUDA fp16 [L, 64]
        view with 16-byte aligned rows (:meth:`_Dev.q16_row_stride`)."""
        if qk_bits not in (8, 4):
            raise ValueError("qk_bits must be 8 or 4")
        if v_bits not in (8, 4) or (v_bits == 4 and qk_bits != 4):
            raise ValueError("v_bits must be 8, or 4 with 4-bit Q/K")
        if p_mode not in (P_REFERENCE, P_ROW_SCALED, P_REFERENCE_4BIT):
            raise ValueError("unknown p_mode")
        self.qk_bits, self.V, self.p_mode, self.p_scale = qk_bits, V, p_mode, float(p_scale)
        self.v_bits = v_bits
        self.fuse_q_requested = bool(fuse_q)
        self.debug = bool(debug)
        # the reference truncates (quantization_kernel.cu:65-69); the row-scaled layer rounds
        self.rounding = rounding if rounding is not None else (ROUND_NEAREST if p_mode == P_ROW_SCALED else ROUND_TRUNC)
        self.v_rounding = self.rounding | (QUANT_RANGE4 if v_range4 and v_bits == 8 else 0)
        self.device = torch.device(device)
        if resident is not None:
            if any(d.qk_bits != qk_bits or d.v_bits != v_bits for d in resident):
                raise ValueError("resident matrices were quantized for another width")
            self.mats = list(resident)
        else:
            if scales is None:
                scales = qkv_scales(mats, qk_bits)
            outs = outputs if outputs is not None else [None] * len(mats)
            if len(outs) != len(mats):
                raise ValueError("one output tensor per matrix")
            self.mats = [make_resident(m, sc, qk_bits, V, self.p_scale, self.device, debug, o, v_bits)
                         for m, sc, o in zip(mats, scales, outs)]
        self._assemble()
        self.graph = None

    @classmethod
    def from_resident(cls, devs: list[_Dev], **kw) -> "AttentionBatch":
        """A batch over matrices whose operands are already allocated (their
        quantization scales, width and output buffers are kept)."""
        if not devs:
            raise ValueError("empty batch")
        return cls([], qk_bits=devs[0].qk_bits, v_bits=devs[0].v_bits, resident=devs, **kw)

    def _assemble(self) -> None:
        """Work list (XCD-aware order, :func:`xcd_item_order`) and descriptors."""
        dev = self.device
        order = xcd_item_order([d.m_vec for d in self.mats], [d.nnz + 16 * d.m_vec for d in self.mats])
        self.n_items = int(len(order)) if any(d.m_vec for d in self.mats) else 0
        self.items = _pinned(order).to(dev, non_blocking=True)
        self.fuse_q = (self.fuse_q_requested and dev.type == "cuda"
                       and all(d.q16_row_stride() is not None for d in self.mats))
        q_rnd = self.rounding if self.fuse_q else None
        self.desc = _bytes_tensor(b"".join(d.desc_bytes(q_rnd) for d in self.mats), dev)
        jobs = [d.quant_job_bytes() for d in self.mats]
        qk_raw, v_raw = b"".join(j[0] for j in jobs), b"".join(j[1] for j in jobs)
        self.qk_jobs = _bytes_tensor(qk_raw, dev)
        self.v_jobs = _bytes_tensor(v_raw, dev)
        # the fused path's first quantize launch: the Q and K jobs, or the K
        # jobs alone (the second 48-byte record of each pair) when the kernel
        # quantizes Q
        n = len(self.mats)
        job = ctypes.sizeof(_QuantJobC)
        first_raw = b"".join(j[0][job:] for j in jobs) if self.fuse_q else qk_raw
        self.n_first = n if self.fuse_q else 2 * n
        self.first_jobs = _bytes_tensor(first_raw, dev) if self.fuse_q else self.qk_jobs
        # V at the width and rounding of Q and K joins their launch (the Q
        # and K jobs, then the V jobs)
        self.all_jobs = (_bytes_tensor(first_raw + v_raw, dev)
                         if self.v_bits == self.qk_bits and self.v_rounding == self.rounding
                         and self.n_first + n <= MAX_QUANT_JOBS else None)
        self.max_count = max((d.L * HEAD_DIM for d in self.mats), default=0)
        # the descriptor copies run on the current stream; launches on any
        # other stream wait for them
        self._ready = torch.cuda.Event() if dev.type == "cuda" else None
        if self._ready is not None:
            self._ready.record()

    def _wait_ready(self, stream) -> None:
        if self._ready is None or torch.cuda.is_current_stream_capturing():
            return
        (stream if stream is not None else torch.cuda.current_stream()).wait_event(self._ready)

    # -- fused path ---------------------------------------------------------
    def quantize(self, stream=None) -> None:
        self._wait_ready(stream)
        lib, st = _lib.lib(), _stream(stream)
        n = len(self.mats)
        if self.all_jobs is not None:
            _lib.check(lib.mc_attn_quantize_batched(_dptr(self.all_jobs), self.n_first + n, self.qk_bits,
                                                    self.rounding, self.max_count, st), "mc_attn_quantize_batched")
            return
        _lib.check(lib.mc_attn_quantize_batched(_dptr(self.first_jobs), self.n_first, self.qk_bits, self.rounding,
                                                self.max_count, st), "mc_attn_quantize_batched")
        _lib.check(lib.mc_attn_quantize_batched(_dptr(self.v_jobs), n, self.v_bits, self.v_rounding,
                                                self.max_count, st), "mc_attn_quantize_batched")

    def attend(self, stream=None, grid_waves: int = DEFAULT_GRID_WAVES) -> None:
        self._wait_ready(stream)
        _lib.check(_lib.lib().mc_attn_forward(_dptr(self.desc), _dptr(self.items), self.n_items, self.qk_bits,
                                              self.v_bits, self.V, self.p_scale, self.p_mode, grid_waves, int(self.debug),
                                              _stream(stream)),
                   "mc_attn_forward")

    def forward(self, stream=None, grid_waves: int = DEFAULT_GRID_WAVES) -> list[torch.Tensor]:
        self.quantize(stream)
        self.attend(stream, grid_waves)
        return [d.o for d in self.mats]

    def capture(self, stream: torch.cuda.Stream, grid_waves: int = DEFAULT_GRID_WAVES) -> torch.cuda.CUDAGraph:
        """Captures quantize + fused attention into a hipGraph on ``stream``."""
        with torch.cuda.stream(stream):
            self.forward(stream, grid_waves)      # warm-up outside the graph
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            self.forward(stream, grid_waves)
        self.graph = g
        return g

    def replay(self, stream: torch.cuda.Stream) -> None:
        with torch.cuda.stream(stream):
            self.graph.replay()

    # -- plain one-op-per-launch path -------------------------------------
    def _plain_buffers(self, d: _Dev, stream) -> dict:
        """Work buffers of the one-op-per-launch pipeline, created on the
        launch stream so their job-descriptor copies precede the launches."""
        w, V = d.work, self.V
        if w:
            return w
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            w["s"] = torch.zeros(max(d.aligned_num * V, 16), dtype=torch.int32, device=self.device)
            w["p"] = torch.zeros(max(d.aligned_num * V, 16), dtype=torch.uint8, device=self.device)
            w["inv_l"] = torch.empty(max(d.m_vec * V, 1), dtype=torch.float32, device=self.device)
            w["o_int"] = torch.empty((d.L, HEAD_DIM), dtype=torch.int32, device=self.device)
            # fp32 result of the pipeline when the matrix's output is fp16
            w["o32"] = (d.o if d.o.dtype == torch.float32 else
                        torch.empty((d.L, HEAD_DIM), dtype=torch.float32, device=self.device))
            # the pipeline's P V runs on int8 V: packed 4-bit V is quantized
            # again into bytes clamped to [-8, 7] (the same integers)
            if d.v_bits == 4:
                w["vq8"] = torch.empty(max(d.L * HEAD_DIM, 16), dtype=torch.uint8, device=self.device)
            vq8, v_rnd = (w["vq8"], self.rounding | QUANT_RANGE4) if d.v_bits == 4 else (d.vq, self.v_rounding)
            for name, (x, s, y, b, rnd) in {"jq": (d.q16, d.scales[0], d.qq, self.qk_bits, self.rounding),
                                            "jk": (d.k16, d.scales[1], d.kq, self.qk_bits, self.rounding),
                                            "jv": (d.v16, d.scales[2], vq8, 8, v_rnd)}.items():
                w[name] = (_struct_tensor([_QuantJobC(x.data_ptr(), x.numel(), s, y.data_ptr(), *_rows(x))],
                                          self.device), b, rnd)
        return w

    def forward_plain(self, stream=None) -> list[torch.Tensor]:
        self._wait_ready(stream)
        lib, st = _lib.lib(), _stream(stream)
        V = self.V
        for d in self.mats:
            w = self._plain_buffers(d, stream)
            for name in ("jq", "jk", "jv"):
                job, b, rnd = w[name]
                _lib.check(lib.mc_attn_quantize_batched(_dptr(job), 1, b, rnd, d.L * HEAD_DIM, st), "quantize")
            if d.m_vec == 0:
                continue
            _lib.check(lib.mc_sddmm_run(self.qk_bits, SIGNED, 1, d.m_vec, HEAD_DIM, d.L, _dptr(d.row_indices),
                                        _dptr(d.row_offsets), _dptr(d.col_indices), _dptr(d.qq), _dptr(d.kq),
                                        _dptr(w["s"]), V, st), "sddmm")
            _lib.check(lib.mc_attn_softmax_requant(d.m_vec, V, _dptr(d.row_indices), _dptr(d.row_offsets),
                                                   _dptr(w["s"]), d.alpha, self.p_scale, self.p_mode,
                                                   _dptr(w["p"]), _dptr(w["inv_l"]), st), "softmax")
            _lib.check(lib.mc_spmm_run(8, 8, SIGNED, -1, d.m_vec, V, HEAD_DIM, d.L, _dptr(d.row_indices),
                                       _dptr(d.row_offsets), _dptr(d.col_indices), _dptr(w["p"]),
                                       _dptr(w.get("vq8", d.vq)), _dptr(w["o_int"]), st), "spmm")
            row_scale = _dptr(w["inv_l"]) if self.p_mode == P_ROW_SCALED else None
            _lib.check(lib.mc_attn_dequant(_dptr(w["o_int"]), d.L, HEAD_DIM, d.o_scale, row_scale, _dptr(w["o32"]),
                                           st), "dequant")
            if d.o.dtype != torch.float32:
                with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
                    d.o.copy_(w["o32"])          # fp16 output: same rounding as the fused kernel's store
        return [d.work["o32"] if d.work else d.o for d in self.mats]

    def hbm_bytes(self) -> int:
        total = 0
        for d in self.mats:
            for t in (d.row_offsets, d.col_indices, d.row_indices, d.q16, d.k16, d.v16, d.qq, d.kq, d.vq, d.o):
                total += t.numel() * t.element_size()
        return total + self.desc.numel() + self.items.numel() * 4


def _np_dtype(struct) -> np.dtype:
    """numpy record dtype with the C struct's field offsets and size, marked
    as an aligned struct so that numpy keeps the padding when it promotes the
    dtype (np.concatenate of unaligned record arrays repacks them)."""
    kinds = {ctypes.c_int: np.int32, ctypes.c_float: np.float32, ctypes.c_void_p: np.uint64,
             ctypes.c_longlong: np.int64}
    names = [n for n, _ in struct._fields_]
    return np.dtype({"names": names, "formats": [kinds[t] for _, t in struct._fields_],
                     "offsets": [getattr(struct, n).offset for n in names], "itemsize": ctypes.sizeof(struct)},
                    align=True)


_ATTN_DT = _np_dtype(_AttnMatrixC)
_JOB_DT = _np_dtype(_QuantJobC)


def _align256(x):
    return (x + 255) // 256 * 256


class ResidentBlock:
    """Matrices of one Q/K width (typically one request) whose quantized
    operands and fp32 outputs share ONE device arena, with their launch
    descriptors, quantize jobs and work-list items built once, vectorized,
    when the block is created.  Blocks are concatenated into a launch by
    :class:`BlockBatch` without per-matrix device work."""

    def __init__(self, mats: list[AttentionMatrix], scales: list[tuple], qk_bits: int, V: int, p_scale: float,
                 device="cuda"):
        if qk_bits not in (8, 4):
            raise ValueError("qk_bits must be 8 or 4")
        dev = torch.device(device)
        self.qk_bits, self.V, self.n = qk_bits, V, len(mats)
        masks = [m.mask if isinstance(m.mask, ResidentMask) else ResidentMask.from_csr(m.mask, dev) for m in mats]
        ins = [(m.q.contiguous(), m.k.contiguous(), m.v.contiguous()) for m in mats]
        for m, mk, (q, _, _) in zip(mats, masks, ins):
            if q.shape[1] != HEAD_DIM or mk.V != V or mk.m_vec * V != m.L:
                raise ValueError("matrix shapes do not match the block")
        self._inputs, self._masks = ins, masks            # kept alive with the block
        n = self.n
        Ls = np.array([m.L for m in mats], dtype=np.int64)
        qk_sz = np.maximum(Ls * (HEAD_DIM * qk_bits // 8), 16)
        v_sz = np.maximum(Ls * HEAD_DIM, 16)
        o_sz = Ls * HEAD_DIM * 4
        per = 2 * _align256(qk_sz) + _align256(v_sz) + _align256(o_sz)
        off = np.cumsum(per) - per
        self.arena_bytes = int(per.sum())
        self.arena = torch.empty(max(self.arena_bytes, 256), dtype=torch.uint8, device=dev)
        base = np.uint64(self.arena.data_ptr())
        q_off = off
        k_off = q_off + _align256(qk_sz)
        v_off = k_off + _align256(qk_sz)
        self._o_off, self._Ls = v_off + _align256(v_sz), Ls
        sc = np.asarray(scales, dtype=np.float32).reshape(n, 3)
        alpha = np.float32(1.0) / (sc[:, 0] * sc[:, 1] * np.float32(math.sqrt(HEAD_DIM)))
        o_scale = np.float32(1.0) / (sc[:, 2] * np.float32(p_scale))
        d = np.zeros(n, dtype=_ATTN_DT)
        d["L"] = Ls
        d["m_vec"] = [mk.m_vec for mk in masks]
        d["row_offsets"] = [mk.row_offsets.data_ptr() for mk in masks]
        d["col_indices"] = [mk.col_indices.data_ptr() for mk in masks]
        d["q"], d["k"], d["v"] = base + q_off.astype(np.uint64), base + k_off.astype(np.uint64), \
            base + v_off.astype(np.uint64)
        d["alpha"], d["o_scale"] = alpha, o_scale
        d["o"] = base + self._o_off.astype(np.uint64)
        self.desc = d
        x_ptr = np.array([[t.data_ptr() for t in trio] for trio in ins], dtype=np.uint64).reshape(n, 3)
        qk = np.zeros((n, 2), dtype=_JOB_DT)
        qk["x"] = x_ptr[:, :2]
        qk["count"] = (Ls * HEAD_DIM)[:, None]
        qk["scale"] = sc[:, :2]
        qk["out"] = np.stack([d["q"], d["k"]], axis=1)
        self.qk_jobs = qk.reshape(-1)
        vj = np.zeros(n, dtype=_JOB_DT)
        vj["x"], vj["count"], vj["scale"], vj["out"] = x_ptr[:, 2], Ls * HEAD_DIM, sc[:, 2], d["v"]
        self.v_jobs = vj
        self.max_count = int(Ls.max()) * HEAD_DIM if n else 0
        self.nnz_vec = np.array([mk.nnz_vec for mk in masks], dtype=np.int64)
        m_vecs = d["m_vec"].astype(np.int64)
        self.n_items = int(m_vecs.sum())
        # per-XCD row lists with block-local matrix indices, built once
        self.xcd_lists = xcd_bins(m_vecs, self.nnz_vec + 16 * m_vecs)

    def outputs(self) -> list[torch.Tensor]:
        """fp32 [L, 64] output view of every matrix."""
        o = self.arena
        return [o[int(a):int(a) + int(L) * HEAD_DIM * 4].view(torch.float32).view(int(L), HEAD_DIM)
                for a, L in zip(self._o_off, self._Ls)]


class BlockBatch:
    """One fused launch (batched quantize + persistent attention) over the
    union of resident blocks of one Q/K width: descriptors concatenated,
    items merged longest-row-first on the device."""

    def __init__(self, blocks: list[ResidentBlock], p_mode: int = P_ROW_SCALED, p_scale: float = 127.0,
                 rounding: int | None = None, device="cuda"):
        if not blocks or any(b.qk_bits != blocks[0].qk_bits or b.V != blocks[0].V for b in blocks):
            raise ValueError("a block batch needs blocks of one width and V")
        dev = torch.device(device)
        self.qk_bits, self.V, self.p_mode, self.p_scale = blocks[0].qk_bits, blocks[0].V, p_mode, float(p_scale)
        self.rounding = rounding if rounding is not None else (ROUND_NEAREST if p_mode == P_ROW_SCALED else ROUND_TRUNC)
        self.blocks = blocks
        self.n_mats = sum(b.n for b in blocks)
        # raw bytes per block: np.concatenate would repack the padded C
        # record layout (numpy drops the padding of structured dtypes when it
        # promotes them)
        self.desc = _bytes_tensor(b"".join(b.desc.tobytes() for b in blocks), dev)
        qk_raw = b"".join(b.qk_jobs.tobytes() for b in blocks)
        v_raw = b"".join(b.v_jobs.tobytes() for b in blocks)
        self.qk_jobs = _bytes_tensor(qk_raw, dev)
        self.v_jobs = _bytes_tensor(v_raw, dev)
        # int8 V joins the Q and K launch when they are int8 too
        self.all_jobs = (_bytes_tensor(qk_raw + v_raw, dev)
                         if self.qk_bits == 8 and 3 * self.n_mats <= MAX_QUANT_JOBS else None)
        self.max_count = max(b.max_count for b in blocks)
        # block j's XCD lists rotate by j so the first (largest) bins of the
        # requests spread over the XCDs; matrix indices shift to the batch's
        base = np.cumsum([0] + [b.n for b in blocks[:-1]])
        lists = []
        for x in range(N_XCD):
            parts = []
            for j, (b0, blk) in enumerate(zip(base, blocks)):
                li = blk.xcd_lists[(x - j) % N_XCD]
                if len(li):
                    parts.append(li + np.array([b0, 0], dtype=np.int32))
            lists.append(np.concatenate(parts) if parts else np.zeros((0, 2), np.int32))
        order = interleave_bins(lists)
        self.n_items = int(len(order)) if sum(b.n_items for b in blocks) else 0
        self.items = _pinned(order).to(dev, non_blocking=True)
        # the descriptor copies run on the current stream; launches on any
        # other stream wait for them
        self._ready = torch.cuda.Event() if dev.type == "cuda" else None
        if self._ready is not None:
            self._ready.record()

    def forward(self, stream=None, grid_waves: int = DEFAULT_GRID_WAVES) -> None:
        if self._ready is not None and not torch.cuda.is_current_stream_capturing():
            (stream if stream is not None else torch.cuda.current_stream()).wait_event(self._ready)
        lib, st = _lib.lib(), _stream(stream)
        if self.all_jobs is not None:
            _lib.check(lib.mc_attn_quantize_batched(_dptr(self.all_jobs), 3 * self.n_mats, 8, self.rounding,
                                                    self.max_count, st), "mc_attn_quantize_batched")
        else:
            _lib.check(lib.mc_attn_quantize_batched(_dptr(self.qk_jobs), 2 * self.n_mats, self.qk_bits,
                                                    self.rounding, self.max_count, st), "mc_attn_quantize_batched")
            _lib.check(lib.mc_attn_quantize_batched(_dptr(self.v_jobs), self.n_mats, 8, self.rounding,
                                                    self.max_count, st), "mc_attn_quantize_batched")
        _lib.check(lib.mc_attn_forward(_dptr(self.desc), _dptr(self.items), self.n_items, self.qk_bits, 8, self.V,
                                       self.p_scale, self.p_mode, grid_waves, 0, st), "mc_attn_forward")


# ---------------------------------------------------------------------------
# Single ops on device tensors: the quantizer (src/cuda/quantization.cpp)
# and the softmax + requantization of int32 logits.
# ---------------------------------------------------------------------------

def bquantization(x: torch.Tensor, bits: int, scale: float, rounding: int = ROUND_TRUNC,
                  stream=None) -> torch.Tensor:
    """fp16 [..., L, D] -> int32-packed [..., L, D*bits/32] (trunc(clamp(x*scale)) by default).
    D * bits must be a multiple of 32 (the reference passes D = 64).

    Runs on ``stream`` (default: the current stream), which must see ``x``
    ready.  The contiguous copy of ``x``, the job descriptor and the output
    are made on that stream too, so the caching allocator cannot hand their
    memory to other work before the launch has read or written it."""
    if bits not in (8, 4):
        raise ValueError("bits must be 8 or 4")
    if x.dim() == 0 or (x.shape[-1] * bits) % 32 != 0:
        raise ValueError(f"the last dimension ({x.shape[-1] if x.dim() else 0}) times bits must be a multiple of 32")
    if not x.is_cuda or x.dtype != torch.float16:
        raise ValueError("x must be a float16 device tensor")
    ctx = torch.cuda.stream(stream) if isinstance(stream, torch.cuda.Stream) else contextlib.nullcontext()
    with ctx:
        x = x.contiguous()
        out = torch.empty(x.shape[:-1] + (x.shape[-1] * bits // 32,), dtype=torch.int32, device=x.device)
        job = _struct_tensor([_QuantJobC(x.data_ptr(), x.numel(), float(scale), out.data_ptr())], x.device)
        _lib.check(_lib.lib().mc_attn_quantize_batched(_dptr(job), 1, bits, rounding, x.numel(), _stream(stream)),
                   "bquantization")
    return out


def softmax_requant(row_indices, row_offsets, s_vals, m_vec: int, vec_length: int, alpha: float, p_scale: float,
                    p_mode: int = P_REFERENCE, aligned_num: int | None = None, stream=None):
    """int32 logits [slot][V] -> int8 P in the SpMM A layout (+ 1/sum per row).
    The reference's fp16-logit op with its signature is
    :func:`magicube_amd.sptrans.q_bcsr_softmax`."""
    n = aligned_num if aligned_num is not None else s_vals.numel() // vec_length
    p = torch.zeros(max(n * vec_length, 16), dtype=torch.uint8, device=s_vals.device)
    inv_l = torch.empty(max(m_vec * vec_length, 1), dtype=torch.float32, device=s_vals.device)
    _lib.check(_lib.lib().mc_attn_softmax_requant(m_vec, vec_length, _dptr(row_indices), _dptr(row_offsets),
                                                  _dptr(s_vals), alpha, p_scale, p_mode, _dptr(p), _dptr(inv_l),
                                                  _stream(stream)), "softmax_requant")
    return p, inv_l
