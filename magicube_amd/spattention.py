"""Sparse multi-head attention and the 4-layer sparse encoder of the
reference's end-to-end evaluation, on the fused MI355X attention kernel.

``SpMultiheadAttention`` mirrors ``spMultiheadAttention``
(end2end_eval/sparse_transformer_magicube/spattention.py:233-250): a
``torch.nn.MultiheadAttention`` whose forward also takes the vector-sparse
mask (``row_indices``, ``row_offsets``, ``column_indices`` as
``static_random_mask_aligned`` builds them, verify/static_mask.py:29-76), the
vector length and the operand precisions.  Per (sample, head) it computes,
as the reference does (spattention.py:51-53, 182-221):

    Q, K, V = in-projection of the input, quantized with the fixed scale 36
              (truncation, clamped)
    P       = trunc(min(softmax(Q K^T / (36 * 36) / sqrt(64)) * 32, 127))
              over the mask's nonzeros
    O       = P V / (36 * 32), then the output projection.

All (sample, head) matrices run in one persistent launch (``mc_attn_forward``)
over the shared mask.  The module keeps its staging buffers, quantized
operands, outputs and launch descriptors per (shape, mask), so later calls
only launch kernels and can be captured in a HIP graph.

Precisions (``lhs_pre`` = P bits, ``rhs_pre`` = Q/K/V bits), the reference's
four pairs (spattention.py:194-215):

* 8, 8: int8 Q, K, V and P;
* 16, 8: the same integers as 8, 8.  P = trunc(min(... * 32, 32767)) never
  exceeds 32, so the reference's 16-bit P holds exactly the int8 P values and
  P V is the same exact int32 product (its bspmm_16b8b splits P into byte
  planes);
* 8, 4: int4 Q, K and V (spattention.py:191); V held one 4-bit value per
  byte: the fused kernel also takes it packed as the reference's bspmm_8b4b
  operand ([L][32] bytes, deq_spmm.cpp:50-84, ``AttentionBatch(v_bits=4)``),
  bit for bit the same, but its 16-bit gathers touch the same L1 lines and
  add the widening, 27 % slower on a 128-matrix batch;
* 4, 4: as 8, 4 with P clamped at 7 (the reference's 4-bit P,
  q_softmax_kernel.cu:84-118).

``key_padding_mask`` ([batch, L], True or nonzero = padded key) removes the
padded key columns from each sample's mask, as ``torch.nn.MultiheadAttention``
does; a row whose every key is padded gets 0.  The reference checks only the
mask's shape (spattention.py:171-181) and its encoder passes none
(end_to_end.py:70-73).  ``need_weights=True`` also returns the head-averaged
weights [batch, L, L]: the layer's quantized probabilities P / 32 at the
mask's nonzeros, 0 elsewhere (from the fused kernel's P output).  The
reference's own ``need_weights`` branch views its packed P as a dense
[batch, heads, L, L] tensor (spattention.py:225-228), which fails for any
sparse mask; its encoder calls with ``need_weights=False``.

``SparseTransformerEncoder`` / ``TransformerEncoder`` restate the encoder
benchmark models (end_to_end.py:47-183): embedding, sinusoidal positions,
four blocks of LayerNorm -> attention -> residual -> LayerNorm -> Linear ->
Linear -> residual, then LayerNorm and a Linear head.
"""
from __future__ import annotations

import math
from collections import OrderedDict

import numpy as np
import torch
import torch.nn.functional as F
from torch import nn

from . import _lib, formats
from .attention import (HEAD_DIM, P_REFERENCE, P_REFERENCE_4BIT, ROUND_TRUNC, AttentionBatch, AttentionMatrix,
                        ResidentMask)

# plans (buffers, quantized operands, descriptors) kept per module, least
# recently used dropped first
MAX_PLANS = 8

SCALE_QKV = 36.0    # spattention.py:51
SCALE_SFMX = 32.0   # spattention.py:53


def static_random_mask_aligned(m: int, n: int, sparsity: float, mma_k_dim: int, seed: int | None = None,
                               device="cuda"):
    """The reference's attention mask builder (verify/static_mask.py:29-76):
    an element-random m x n pattern of density 1 - sparsity (each vector row
    of the attention matrix is one row here), row pairs padded to
    ``mma_k_dim`` slots, -1-padded column indices, the same indices shuffled
    in groups of 8, and the rows ordered by ascending length.  Returns
    (column_indices, column_indices_shuffle, row_offsets, row_indices,
    aligned_num_item) with the tensors on ``device``."""
    rng = np.random.default_rng(seed)
    k = int(round((1.0 - sparsity) * m * n))
    flat = np.sort(rng.choice(m * n, size=k, replace=False))
    rows, cols = flat // n, (flat % n).astype(np.int32)
    counts = np.bincount(rows, minlength=m)
    ro = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    pairs, an = formats.align_rows(ro, mma_k_dim)
    aligned = formats.align_cols(ro, cols, pairs, an, shuffle=False)
    shuffled = formats.align_cols(ro, cols, pairs, an, shuffle=True)
    order = np.argsort(counts, kind="stable").astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(device)  # noqa: E731
    return t(aligned), t(shuffled), t(pairs), t(order), int(an)


def mask_from_reference(row_indices: torch.Tensor, row_offsets: torch.Tensor, column_indices: torch.Tensor,
                        seq_len: int, vec_length: int) -> ResidentMask:
    """A :class:`ResidentMask` over the reference's device mask tensors
    (aligned (start, end) pairs, -1-padded unshuffled column indices).  The
    tensors are used in place; row starts must be multiples of 16 slots
    (alignment 16 or 32, as the reference pads for 8- and 4-bit operands)."""
    pairs = row_offsets.detach().to("cpu", torch.int32).numpy()
    m_vec = pairs.size // 2
    if m_vec * vec_length != seq_len:
        raise ValueError(f"mask has {m_vec} vector rows of {vec_length}, sequence length is {seq_len}")
    starts, ends = pairs[0::2], pairs[1::2]
    if (starts % 16).any() or (ends < starts).any() or (ends.max(initial=0) > column_indices.numel()):
        raise ValueError("row_offsets must hold (start, end) pairs with starts aligned to 16 slots")
    ro = row_offsets.to(torch.int32).contiguous()
    return ResidentMask(L=seq_len, V=vec_length, m_vec=m_vec, nnz_vec=int((ends - starts).sum()),
                        aligned_num=int(column_indices.numel()), row_offsets=ro,
                        col_indices=column_indices.to(torch.int32).contiguous(),
                        row_indices=row_indices.to(torch.int32).contiguous(),
                        row_nnz=(ro[1::2] - ro[0::2]).contiguous(), pairs=pairs)


def _real_slots(m: ResidentMask):
    """(slot, vector row) of every real nonzero of a mask, in row order."""
    pairs = m.pairs if m.pairs is not None else m.row_offsets.detach().to("cpu").numpy()
    starts, ends = pairs[0::2].astype(np.int64), pairs[1::2].astype(np.int64)
    counts = ends - starts
    rows = np.repeat(np.arange(m.m_vec), counts)
    slots = np.repeat(starts - np.concatenate([[0], np.cumsum(counts)[:-1]]), counts) + np.arange(counts.sum())
    return slots, rows


def _padded_mask(base: ResidentMask, base_cols: np.ndarray, drop: np.ndarray, device) -> ResidentMask:
    """``base`` without the key columns where ``drop`` is True, rows
    re-aligned to 16 slots."""
    slots, rows = _real_slots(base)
    cols = base_cols[slots]
    keep = ~drop[cols]
    ro = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=base.m_vec))]).astype(np.int32)
    csr = formats.VectorCSR(base.m_vec, base.L, base.V, ro, cols[keep].astype(np.int32))
    return ResidentMask.from_csr(csr, device)


class _Plan:
    """Per (shape, mask, padding, precision) state of SpMultiheadAttention.

    The in-projection writes ``qkv`` ([L][batch][3 * E] fp16); the quantize
    launch reads each (sample, head)'s [L][64] slices of it in place, and the
    fused kernel writes each head's output as fp16 straight into ``o``
    ([L][batch][E], the output projection's input), so no permute copies run
    between the projections and the attention kernels.  ``masks`` holds one
    mask per sample (the same one without a padding mask).  ``keep`` holds
    the caller's tensors the plan was keyed on, so their addresses cannot be
    handed to other tensors while the plan exists."""

    def __init__(self, L: int, bsz: int, heads: int, masks: list, lhs_pre: int, rhs_pre: int, device, keep=()):
        E = heads * HEAD_DIM
        self.L, self.bsz, self.heads, self.masks, self.keep = L, bsz, heads, masks, tuple(keep)
        self.lhs_pre, self.rhs_pre, self.device = lhs_pre, rhs_pre, device
        self.qkv = torch.empty((L, bsz, 3 * E), dtype=torch.float16, device=device)
        self.o = torch.empty((L, bsz, E), dtype=torch.float16, device=device)
        self.batch = self._batch(debug=False)
        self.debug_batch = None
        self.weight_index = None

    def _batch(self, debug: bool) -> AttentionBatch:
        L, bsz, heads, E = self.L, self.bsz, self.heads, self.heads * HEAD_DIM
        q5 = self.qkv.view(L, bsz, 3, heads, HEAD_DIM)
        o4 = self.o.view(L, bsz, heads, HEAD_DIM)
        mats, outs = [], []
        for b in range(bsz):
            for h in range(heads):
                mats.append(AttentionMatrix(q5[:, b, 0, h], q5[:, b, 1, h], q5[:, b, 2, h], self.masks[b]))
                outs.append(o4[:, b, h])
        n_mat = bsz * heads
        return AttentionBatch(mats, qk_bits=self.rhs_pre, V=self.masks[0].V,
                              p_mode=P_REFERENCE_4BIT if self.lhs_pre == 4 else P_REFERENCE, p_scale=SCALE_SFMX,
                              scales=[(SCALE_QKV, SCALE_QKV, SCALE_QKV)] * n_mat, rounding=ROUND_TRUNC,
                              device=self.device, outputs=outs, v_range4=self.rhs_pre == 4, debug=debug)

    def forward_with_weights(self) -> torch.Tensor:
        """Runs the layer through the kernel that also writes P and returns
        the head-averaged weights P / 32 as [batch, L, L] (float32)."""
        if self.debug_batch is None:
            self.debug_batch = self._batch(debug=True)
            self.weight_index = []
            for m in self.masks:
                # every real (slot, v) of the mask: its byte in the P layout
                # [(slot / 16)][v][16] and its (row, column) in the dense weights
                slots, rows = _real_slots(m)
                V = m.V
                v = np.arange(V)
                p_idx = ((slots[:, None] // 16) * 16 * V + 16 * v[None, :] + slots[:, None] % 16).ravel()
                r_idx = (rows[:, None] * V + v[None, :]).ravel()
                t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)  # noqa: E731
                c_idx = m.col_indices[t(slots)].long().repeat_interleave(V)
                self.weight_index.append((t(p_idx), t(r_idx), c_idx))
        self.debug_batch.forward()
        w = torch.zeros((self.bsz, self.L, self.L), dtype=torch.float32, device=self.device)
        mats = self.debug_batch.mats
        for b, (p_idx, r_idx, cols) in enumerate(self.weight_index):
            acc = torch.zeros(p_idx.numel(), dtype=torch.float32, device=self.device)
            for h in range(self.heads):
                acc += mats[b * self.heads + h].p_out[p_idx].to(torch.int8).float()
            w[b].index_put_((r_idx, cols), acc / (SCALE_SFMX * self.heads))
        return w


def _key_tensors(*ts):
    """Identity of the caller's mask tensors: address, size and version
    counter (in-place writes bump it)."""
    return tuple((t.data_ptr(), t.numel(), t._version) if t is not None else None for t in ts)


class SpMultiheadAttention(nn.MultiheadAttention):
    """Drop-in for the reference's ``spMultiheadAttention`` (sequence-first
    [L, batch, embed] inputs, head dimension 64)."""

    def __init__(self, embed_dim, num_heads, dropout=0.0, bias=True, add_bias_kv=False, add_zero_attn=False,
                 kdim=None, vdim=None):
        super().__init__(embed_dim, num_heads, dropout, bias, add_bias_kv, add_zero_attn, kdim, vdim)
        if embed_dim // num_heads != HEAD_DIM:
            raise ValueError(f"head dimension must be {HEAD_DIM}")
        if add_bias_kv or add_zero_attn:
            raise ValueError("bias_kv and zero attention are not supported")
        self._plans: OrderedDict = OrderedDict()

    def clear_plans(self) -> None:
        """Drops every cached plan (and the device memory it holds)."""
        self._plans.clear()

    def forward(self, query, key, value, key_padding_mask=None, need_weights=True, row_indices=None,
                row_offsets=None, column_indices=None, vec_length=2, lhs_pre=8, rhs_pre=8, mask=None):
        if (lhs_pre, rhs_pre) not in ((8, 8), (16, 8), (8, 4), (4, 4)):
            raise NotImplementedError(f"precisions lhs_pre={lhs_pre}, rhs_pre={rhs_pre}: the reference's pairs "
                                      "are (8, 8), (16, 8), (8, 4) and (4, 4)")
        p_bits = 4 if lhs_pre == 4 else 8       # 16-bit P holds the int8 P values (P <= 32)
        L, bsz, E = query.shape
        H = self.num_heads
        if key.shape != query.shape or value.shape != query.shape:
            raise ValueError("query, key and value must have the same shape")
        if key_padding_mask is not None and tuple(key_padding_mask.shape) != (bsz, L):
            raise ValueError(f"key_padding_mask must be [batch, L] = [{bsz}, {L}]")
        if mask is None:
            if row_offsets is None or column_indices is None:
                raise ValueError("a mask (row_offsets, column_indices) is required")
            key_ = (L, bsz, p_bits, rhs_pre, vec_length) + _key_tensors(row_offsets, column_indices, row_indices)
        else:
            key_ = (L, bsz, p_bits, rhs_pre, mask.V, id(mask))
        if key_padding_mask is not None:
            key_ += _key_tensors(key_padding_mask)
        plan = self._plans.get(key_)
        if plan is None:
            keep = [t for t in (row_offsets, column_indices, row_indices, key_padding_mask) if t is not None]
            if mask is None:
                ri = row_indices if row_indices is not None else torch.zeros(1, dtype=torch.int32,
                                                                               device=query.device)
                mask = mask_from_reference(ri, row_offsets, column_indices, L, vec_length)
            else:
                keep.append(mask)
            if key_padding_mask is None:
                masks = [mask] * bsz
            else:
                drop = key_padding_mask.detach().to("cpu").numpy().astype(bool)
                base_cols = mask.col_indices.detach().to("cpu").numpy()
                uniq = {}
                masks = []
                for b in range(bsz):
                    k = drop[b].tobytes()
                    if k not in uniq:
                        uniq[k] = _padded_mask(mask, base_cols, drop[b], query.device)
                    masks.append(uniq[k])
            plan = _Plan(L, bsz, H, masks, p_bits, rhs_pre, query.device, keep)
            self._plans[key_] = plan
            while len(self._plans) > MAX_PLANS:
                self._plans.popitem(last=False)
        else:
            self._plans.move_to_end(key_)

        # in-projection into the plan's [L][batch][3E] buffer, read in place
        # by the quantize launch (K, V) and the fused kernel (Q)
        if query.dtype != torch.float16:
            raise NotImplementedError("the quantized path takes fp16 activations")
        w, b = self.in_proj_weight, self.in_proj_bias
        qkv2 = plan.qkv.view(L * bsz, 3 * E)
        if query is key and key is value:
            x2 = query.reshape(L * bsz, E)
            if b is None:
                torch.mm(x2, w.t(), out=qkv2)
            else:
                torch.addmm(b, x2, w.t(), out=qkv2)
        else:
            for i, x in enumerate((query, key, value)):
                part = F.linear(x.reshape(L * bsz, E), w[i * E:(i + 1) * E], None if b is None else b[i * E:(i + 1) * E])
                qkv2[:, i * E:(i + 1) * E].copy_(part)
        if need_weights:
            weights = plan.forward_with_weights().to(query.dtype)
        else:
            plan.batch.forward()
            weights = None
        return F.linear(plan.o, self.out_proj.weight, self.out_proj.bias), weights


class FusedLayerNorm(nn.LayerNorm):
    """nn.LayerNorm over the last dimension whose fp16 inference forward is
    one HIP launch (``mc_layernorm_f16``), with the residual add before it
    fused when ``residual`` is given: LN(half(x + residual)).  Other inputs
    (other dtypes or widths, autograd) take torch's own add and LayerNorm,
    which the reference's encoder uses (end_to_end.py:77-89)."""

    def _fast(self, x, residual) -> bool:
        """The HIP kernel's preconditions: fp16 operands on the current
        device, contiguous, x and residual 8-byte aligned (four halves per
        load), a supported width, and no operand that needs a gradient.
        Checked with cheap attribute reads: the encoder's eager forward is
        bound by host time."""
        n = self.normalized_shape[0] if len(self.normalized_shape) == 1 else 0
        if (n not in (256, 512, 1024) or not x.is_cuda or x.dtype != torch.float16 or x.shape[-1] != n
                or not x.is_contiguous() or x.data_ptr() % 8):
            return False
        dev = x.get_device()
        if dev != torch.cuda.current_device():
            return False
        grad = torch.is_grad_enabled()
        if grad and x.requires_grad:
            return False
        if residual is not None and (residual.shape != x.shape or residual.data_ptr() % 8):
            return False
        for t in (residual, self.weight, self.bias):
            if t is not None and (t.dtype != torch.float16 or t.get_device() != dev or not t.is_contiguous()
                                  or (grad and t.requires_grad)):
                return False
        return True

    def forward(self, x, residual=None):
        w, b = self.weight, self.bias
        if not self._fast(x, residual):
            return super().forward(x if residual is None else x + residual)
        out = torch.empty_like(x)
        # plain integer addresses: the bound argument types are void pointers
        st = _lib.lib().mc_layernorm_f16(x.numel() // x.shape[-1], x.shape[-1], x.data_ptr(),
                                         None if residual is None else residual.data_ptr(),
                                         None if w is None else w.data_ptr(), None if b is None else b.data_ptr(),
                                         float(self.eps), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        if st:
            _lib.check(st, "mc_layernorm_f16")
        return out


class PositionalEncoding(nn.Module):
    """Sinusoidal positions added to [L, batch, embed] (end_to_end.py:115-130)."""

    def __init__(self, d_model: int, max_len: int = 10000):
        super().__init__()
        position = torch.arange(0, max_len, dtype=torch.float).unsqueeze(1)
        div_term = torch.exp(torch.arange(0, d_model, 2).float() * (-math.log(10000.0) / d_model))
        pe = torch.zeros(max_len, d_model)
        pe[:, 0::2] = torch.sin(position * div_term)
        pe[:, 1::2] = torch.cos(position * div_term)
        self.register_buffer("pe", pe.unsqueeze(1))

    def forward(self, x):
        return x + self.pe[:x.size(0)]


class SparseTransformerBlock(nn.Module):
    """end_to_end.py:47-89: LayerNorm -> sparse attention -> residual ->
    LayerNorm -> Linear -> Linear -> residual, with its own random mask."""

    def __init__(self, embed_dim, mlp_dim, num_heads, sparsity, seq_len, vec_length=8, lhs_pre=8, rhs_pre=8,
                 seed=None, device="cuda"):
        super().__init__()
        self.vec_length, self.lhs_pre, self.rhs_pre = vec_length, lhs_pre, rhs_pre
        self.self_attention = SpMultiheadAttention(embed_dim=embed_dim, num_heads=num_heads, dropout=0)
        mma_k_dim = 32 if rhs_pre == 4 else 16
        (self.column_indices, self.column_indices_shuffle, self.row_offsets, self.row_indices,
         self.aligned_num_item) = static_random_mask_aligned(seq_len // vec_length, seq_len, sparsity, mma_k_dim,
                                                             seed=seed, device=device)
        self.layer_norm1 = FusedLayerNorm(normalized_shape=embed_dim)
        self.layer_norm2 = FusedLayerNorm(normalized_shape=embed_dim)
        self.linear1 = nn.Linear(embed_dim, mlp_dim)
        self.linear2 = nn.Linear(mlp_dim, embed_dim)

    def forward(self, x):
        out = self.layer_norm1(x)
        out = self.self_attention(out, out, out, need_weights=False, row_indices=self.row_indices,
                                  row_offsets=self.row_offsets, column_indices=self.column_indices,
                                  vec_length=self.vec_length, lhs_pre=self.lhs_pre, rhs_pre=self.rhs_pre)
        out = self.layer_norm2(out[0], residual=x)      # LN(out + x), the add fused
        out = self.linear1(out)
        out = self.linear2(out)
        return out + x


class TransformerBlock(nn.Module):
    """The dense block of the same benchmark (end_to_end.py:91-113), with
    torch's multi-head attention."""

    def __init__(self, embed_dim, mlp_dim, num_heads):
        super().__init__()
        self.self_attention = nn.MultiheadAttention(embed_dim=embed_dim, num_heads=num_heads, dropout=0)
        self.layer_norm1 = FusedLayerNorm(normalized_shape=embed_dim)
        self.layer_norm2 = FusedLayerNorm(normalized_shape=embed_dim)
        self.linear1 = nn.Linear(embed_dim, mlp_dim)
        self.linear2 = nn.Linear(mlp_dim, embed_dim)

    def forward(self, x):
        out = self.layer_norm1(x)
        out = self.self_attention(out, out, out, need_weights=False)
        out = self.layer_norm2(out[0], residual=x)
        out = self.linear1(out)
        out = self.linear2(out)
        return out + x


class _Encoder(nn.Module):
    def __init__(self, vocab_size, embed_dim, num_class, blocks):
        super().__init__()
        self.encoder = nn.Embedding(num_embeddings=vocab_size, embedding_dim=embed_dim)
        self.pos_encoder = PositionalEncoding(embed_dim)
        self.blocks = nn.ModuleList(blocks)
        self.embed_dim = embed_dim
        self.layer_norm = FusedLayerNorm(normalized_shape=embed_dim)
        self.linear = nn.Linear(embed_dim, num_class)

    def forward(self, x):
        # the embedding of the transposed tokens: the reference's
        # transpose(embedding(x), 0, 1) value for value, but laid out
        # contiguously as [L][batch][E], so that the first block's LayerNorms
        # take the fused kernel
        out = self.encoder(x.t()) * math.sqrt(self.embed_dim)
        out = self.pos_encoder(out)
        for blk in self.blocks:
            out = blk(out)
        out = self.layer_norm(out)
        return self.linear(out)


class SparseTransformerEncoder(_Encoder):
    """end_to_end.py:133-157 (four sparse blocks)."""

    def __init__(self, vocab_size, embed_dim, num_heads, sparsity, seq_len, mlp_dim, num_class, vec_length=8,
                 lhs_pre=8, rhs_pre=8, seed=0, device="cuda", layers=4):
        blocks = [SparseTransformerBlock(embed_dim, mlp_dim, num_heads, sparsity, seq_len, vec_length, lhs_pre,
                                         rhs_pre, seed=None if seed is None else seed + i, device=device)
                  for i in range(layers)]
        super().__init__(vocab_size, embed_dim, num_class, blocks)


class TransformerEncoder(_Encoder):
    """end_to_end.py:160-183 (four dense blocks)."""

    def __init__(self, vocab_size, embed_dim, num_heads, seq_len, mlp_dim, num_class, layers=4):
        super().__init__(vocab_size, embed_dim, num_class,
                         [TransformerBlock(embed_dim, mlp_dim, num_heads) for _ in range(layers)])
