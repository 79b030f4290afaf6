"""The C-ABI library loads, exports every symbol the headers declare, and
rejects bad arguments before any device work (CPU-only checks)."""
import ctypes
import os
import re

import pytest
import torch

from magicube_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mc_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_loads(mclib):
    assert mclib.mc_version() >= 1


def test_every_declared_symbol_is_exported(mclib):
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(mclib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_binding():
    names = declared_functions()
    unbound = sorted(n for n in names if n not in _lib.exported_symbols())
    assert not unbound, unbound


def test_status_strings(mclib):
    for s in range(7):
        assert _lib.status_string(s)


NULL = None


def _spmm(mclib, name, **kw):
    args = dict(m_vec=4, V=8, n=64, k=64, ri=NULL, ro=ctypes.c_void_p(16), ci=ctypes.c_void_p(16),
                vals=ctypes.c_void_p(16), rhs=ctypes.c_void_p(16), out=ctypes.c_void_p(16), sign=0, stream=NULL)
    args.update(kw)
    fn = getattr(mclib, name)
    return fn(args["m_vec"], args["V"], args["n"], args["k"], args["ri"], args["ro"], args["ci"], args["vals"],
              args["rhs"], args["out"], args["sign"], args["stream"])


@pytest.mark.parametrize("name", ["mc_spmm_l8r8", "mc_spmm_l16r8", "mc_spmm_l8r4", "mc_spmm_l4r4", "mc_spmm_l12r8",
                                  "mc_spmm_l16r16", "mc_spmm_l12r4", "mc_spmm_l16r4"])
def test_spmm_validation(mclib, name):
    assert _spmm(mclib, name, V=3) == _lib.MC_EINVAL_V
    assert _spmm(mclib, name, V=16) == _lib.MC_EINVAL_V
    assert _spmm(mclib, name, n=0) == _lib.MC_EINVAL_SHAPE
    assert _spmm(mclib, name, n=6 if not name.endswith("r16") else 7) == _lib.MC_EINVAL_SHAPE  # n % (32/bits(B))
    assert _spmm(mclib, name, k=-1) == _lib.MC_EINVAL_SHAPE
    assert _spmm(mclib, name, sign=7) == _lib.MC_EINVAL_ARG
    assert _spmm(mclib, name, ro=NULL) == _lib.MC_EINVAL_ARG
    assert _spmm(mclib, name, out=NULL) == _lib.MC_EINVAL_ARG
    assert _spmm(mclib, name, k=1 << 20, n=1 << 16) == _lib.MC_EINVAL_SHAPE  # B over 1 GiB
    assert _spmm(mclib, name, k=1 << 18, n=1 << 14) == _lib.MC_EINVAL_SHAPE
    # an empty matrix is a no-op that succeeds without launching
    assert _spmm(mclib, name, m_vec=0) == _lib.MC_OK


def test_spmm_run_validation(mclib):
    P = ctypes.c_void_p(16)
    for lhs, rhs in ((12, 16), (8, 16), (4, 8), (4, 16), (32, 8), (8, 2)):
        assert mclib.mc_spmm_run(lhs, rhs, 0, -1, 4, 8, 64, 64, None, P, P, P, P, P, None) == _lib.MC_EUNSUPPORTED
    assert mclib.mc_spmm_run(8, 8, 0, 99, 4, 8, 64, 64, None, P, P, P, P, P, None) == _lib.MC_EINVAL_ARG
    # config 1 gathers two dwords per lane: n/4 must be even
    assert mclib.mc_spmm_run(8, 8, 0, 1, 4, 8, 68, 64, None, P, P, P, P, P, None) == _lib.MC_EINVAL_SHAPE
    # four dwords per lane are built for 8- and 16-bit B only; the B-stationary kernel for 8-bit B
    assert mclib.mc_spmm_run(8, 4, 0, 2, 4, 8, 512, 64, None, P, P, P, P, P, None) == _lib.MC_EUNSUPPORTED
    assert mclib.mc_spmm_run(16, 4, 0, 2, 4, 8, 512, 64, None, P, P, P, P, P, None) == _lib.MC_EUNSUPPORTED
    assert mclib.mc_spmm_run(16, 16, 0, 8, 4, 8, 512, 64, None, P, P, P, P, P, None) == _lib.MC_EUNSUPPORTED


def test_tile_config_table(mclib):
    from magicube_amd import spmm
    cfgs = spmm.tile_configs()
    assert len(cfgs) == mclib.mc_spmm_num_tile_configs() >= 4
    assert all(c[4] in (0, 1) and c[1] >= 1 and c[2] >= 1 and c[3] >= 1 and 0 <= c[5] <= 10 for c in cfgs)
    assert len(cfgs[0]) == 6
    assert all(c[0] in (1, 2, 4) for c in cfgs if c[4] == 0)
    assert all(c[0] in (8, 16) for c in cfgs if c[4] == 1)
    stationary = [i for i, c in enumerate(cfgs) if c[4] == 1]
    assert stationary
    # the entry points keep no tile state: there is no setter to change what
    # the named entry points run
    assert not hasattr(mclib, "mc_spmm_set_tile_config")


def test_format_validation(mclib):
    assert mclib.mc_format_row_alignment(8, 8) == 16
    assert mclib.mc_format_row_alignment(8, 4) == 32
    assert mclib.mc_format_row_alignment(4, 4) == 32
    assert mclib.mc_format_align_rows(-1, None, 16, None) == -1


def test_sddmm_validation(mclib):
    p = ctypes.c_void_p(16)
    for fn in (mclib.mc_sddmm_l8r8, mclib.mc_sddmm_l4r4, mclib.mc_sddmm_l16r16):
        assert fn(4, 100, 64, None, p, p, p, p, p, 8, 0, None) == _lib.MC_EINVAL_SHAPE   # k % 64
        assert fn(4, 64, 64, None, p, p, p, p, p, 3, 0, None) == _lib.MC_EINVAL_V
        assert fn(4, 1 << 12, 1 << 20, None, p, p, p, p, p, 8, 0, None) == _lib.MC_EINVAL_SHAPE  # rhs over 1 GiB
        assert fn(4, 64, 64, None, None, p, p, p, p, 8, 0, None) == _lib.MC_EINVAL_ARG
        assert fn(4, 64, 64, None, p, p, p, p, p, 8, 5, None) == _lib.MC_EINVAL_ARG
        assert fn(0, 64, 64, None, p, p, p, p, p, 8, 0, None) == _lib.MC_OK
    assert mclib.mc_sddmm_run(8, 0, 7, 4, 64, 64, None, p, p, p, p, p, 8, None) == _lib.MC_EINVAL_ARG
    for bits in (12, 2, 32):
        assert mclib.mc_sddmm_run(bits, 0, 0, 4, 64, 64, None, p, p, p, p, p, 8, None) == _lib.MC_EUNSUPPORTED


def test_batched_and_reference_op_validation(mclib):
    """Argument checks of the batched SpMM / SDDMM launches and the
    reference-semantics softmax and dequantization (no launch happens)."""
    P = ctypes.c_void_p(16)
    stationary = [i for i, c in enumerate(__import__("magicube_amd.spmm", fromlist=["x"]).tile_configs())
                  if c[4] == 1]
    # batch bounds, negative strides, B-stationary configs in a batch
    for batch in (0, 65536):
        assert mclib.mc_spmm_batched(8, 8, 0, -1, batch, 4, 8, 64, 64, None, P, P, P, 0, P, 0, P, 0,
                                     None) == _lib.MC_EINVAL_SHAPE
        assert mclib.mc_sddmm_batched(8, 0, 1, batch, 4, 64, 64, None, P, P, P, 0, P, 0, P, 0, 8,
                                      None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_spmm_batched(8, 8, 0, -1, 2, 4, 8, 64, 64, None, P, P, P, -4, P, 0, P, 0,
                                 None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_spmm_batched(8, 8, 0, stationary[0], 2, 4, 8, 512, 64, None, P, P, P, 0, P, 0, P, 0,
                                 None) == _lib.MC_EUNSUPPORTED
    # byte strides that would start a matrix off a dword boundary
    for vs, rs in ((6, 0), (0, 2)):
        assert mclib.mc_spmm_batched(8, 8, 0, -1, 2, 4, 8, 64, 64, None, P, P, P, vs, P, rs, P, 0,
                                     None) == _lib.MC_EINVAL_SHAPE
        assert mclib.mc_sddmm_batched(8, 0, 1, 2, 4, 64, 64, None, P, P, P, vs, P, rs, P, 0, 8,
                                      None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_spmm_batched(8, 8, 0, -1, 2, 0, 8, 64, 64, None, P, P, P, 0, P, 0, P, 0, None) == _lib.MC_OK
    # softmax on fp16 logits: widths, V, batch, alignment of the 16-byte loads
    assert mclib.mc_attn_softmax_f16(1, 4, 8, None, P, P, 0, 0.125, 32.0, 12, P, 0, None) == _lib.MC_EUNSUPPORTED
    assert mclib.mc_attn_softmax_f16(1, 4, 3, None, P, P, 0, 0.125, 32.0, 8, P, 0, None) == _lib.MC_EINVAL_V
    assert mclib.mc_attn_softmax_f16(0, 4, 8, None, P, P, 0, 0.125, 32.0, 8, P, 0, None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_attn_softmax_f16(1, 4, 8, None, P, ctypes.c_void_p(18), 0, 0.125, 32.0, 8, P, 0,
                                     None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_attn_softmax_f16(2, 4, 8, None, P, P, 12, 0.125, 32.0, 8, P, 0, None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_attn_softmax_f16(1, 0, 8, None, P, P, 0, 0.125, 32.0, 8, P, 0, None) == _lib.MC_OK
    assert mclib.mc_attn_dequant_f16(P, -1, 1.0, P, None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_attn_dequant_f16(None, 4, 1.0, P, None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_attn_dequant_f16(P, 0, 1.0, P, None) == _lib.MC_OK


def test_layernorm_validation(mclib):
    P = ctypes.c_void_p(16)
    assert mclib.mc_layernorm_f16(4, 384, P, None, None, None, 1e-5, P, None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_layernorm_f16(-1, 256, P, None, None, None, 1e-5, P, None) == _lib.MC_EINVAL_SHAPE
    assert mclib.mc_layernorm_f16(4, 256, None, None, None, None, 1e-5, P, None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_layernorm_f16(4, 256, ctypes.c_void_p(18), None, None, None, 1e-5, P, None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_layernorm_f16(4, 256, P, ctypes.c_void_p(20), None, None, 1e-5, P, None) == _lib.MC_EINVAL_ARG
    assert mclib.mc_layernorm_f16(0, 256, None, None, None, None, 1e-5, None, None) == _lib.MC_OK


def test_sptrans_rejects_host_tensors():
    """The reference-signature ops take device tensors only (no CPU path)."""
    from magicube_amd import sptrans
    t = torch.zeros(16, dtype=torch.int32)
    with pytest.raises(ValueError):
        sptrans.bsddmm_8b(t, t, t, t.view(1, 4, 4), t.view(1, 4, 4), 2, 8, 1.0)
    with pytest.raises(ValueError):
        sptrans.bspmm_8b(t, t, t, t, t.view(1, 4, 4), 2, 8, 8, 1.0)
    with pytest.raises(ValueError):
        sptrans.q_bcsr_softmax(t, t, torch.zeros(16, dtype=torch.float16), 0.125, 32.0, 2, 1, 8)


def test_spmm_sddmm_mirrors_reject_host_tensors():
    """The ctypes mirrors never hand a host pointer to a launch."""
    from magicube_amd import sddmm, spmm
    t = torch.zeros(64, dtype=torch.int32)
    with pytest.raises(ValueError):
        spmm.wmma_spmm(8, 8, 1, 8, 16, 16, None, t, t, t, t, t)
    with pytest.raises(ValueError):
        sddmm.wmma_sddmm(8, 1, 64, 16, None, t, t, t, t, t, 8)
