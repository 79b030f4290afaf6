"""The encoder's fused residual + LayerNorm (mc_layernorm_f16) against a
plain PyTorch fp32 LayerNorm of the same fp16 input: within one fp16 step
of the fp32 result plus 2^-17 (the kernel rounds its fp32 value to the
nearest fp16; its fp32 statistics and affine terms differ from the
reference's in the last bits, which matters in absolute terms where the
bias cancels the normalised value).  A host emulation of the kernel's
arithmetic stays within half that bound on 8195-row inputs."""
import ctypes

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from magicube_amd import _lib, spattention

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _run(x, res, w, b, eps, out=None):
    out = torch.empty_like(x) if out is None else out
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().mc_layernorm_f16(x.numel() // x.shape[-1], x.shape[-1], _p(x), _p(res), _p(w), _p(b),
                                           eps, _p(out), st), "mc_layernorm_f16")
    torch.cuda.synchronize()
    return out


def _reference(x, res, w, b, eps):
    s = x if res is None else (x + res)                  # fp16 add, rounded as torch stores it
    return F.layer_norm(s.float(), (x.shape[-1],), None if w is None else w.float(),
                        None if b is None else b.float(), eps)


def _within_one_step(got, ref):
    got, ref = got.float().cpu().numpy(), ref.cpu().numpy()
    tol = np.abs(ref) * 2.0 ** -10 + 2.0 ** -17
    bad = np.abs(got - ref) > tol
    assert not bad.any(), f"{int(bad.sum())} values off by more than one fp16 step, worst {np.abs(got - ref).max()}"


@pytest.mark.parametrize("affine", [True, False])
@pytest.mark.parametrize("residual", [True, False])
@pytest.mark.parametrize("rows", [1, 7, 8195])
@pytest.mark.parametrize("cols", [256, 512, 1024])
def test_layernorm_matches_fp32(cols, rows, residual, affine):
    g = torch.Generator(device="cuda").manual_seed(cols + rows)
    x = (torch.randn(rows, cols, device="cuda", generator=g) * 3 + 1).half()
    res = (torch.randn(rows, cols, device="cuda", generator=g) * 2).half() if residual else None
    w = (torch.rand(cols, device="cuda", generator=g) + 0.5).half() if affine else None
    b = torch.randn(cols, device="cuda", generator=g).half() if affine else None
    print(f"layernorm cols={cols} rows={rows} residual={residual} affine={affine}", flush=True)
    _within_one_step(_run(x, res, w, b, 1e-5), _reference(x, res, w, b, 1e-5))


def test_layernorm_in_place_and_constant_rows():
    """out may alias x; a constant row normalises to the bias exactly."""
    x = torch.randn(64, 256, device="cuda").half()
    x[3] = 2.5
    w = torch.ones(256, device="cuda").half()
    b = torch.full((256,), 0.25, device="cuda").half()
    ref = _reference(x, None, w, b, 1e-5)
    got = _run(x, None, w, b, 1e-5, out=x)
    _within_one_step(got, ref)
    assert torch.all(got[3] == 0.25)


def test_fused_layernorm_module():
    """FusedLayerNorm: the HIP path within one step of torch's fp32
    LayerNorm, with and without the residual; fp32 inputs and autograd take
    torch's own LayerNorm unchanged."""
    torch.manual_seed(3)
    ln = spattention.FusedLayerNorm(256).cuda().half()
    with torch.no_grad():
        ln.weight.uniform_(0.5, 1.5)
        ln.bias.uniform_(-0.5, 0.5)
    x = torch.randn(128, 2, 256, device="cuda").half()
    r = torch.randn(128, 2, 256, device="cuda").half()
    with torch.no_grad():
        _within_one_step(ln(x), _reference(x, None, ln.weight, ln.bias, ln.eps))
        _within_one_step(ln(x, residual=r), _reference(x, r, ln.weight, ln.bias, ln.eps))
        ln32 = spattention.FusedLayerNorm(256).cuda()
        x32 = x.float()
        assert torch.equal(ln32(x32), F.layer_norm(x32, (256,), ln32.weight, ln32.bias, ln32.eps))
    xg = x.clone().requires_grad_(True)
    y = ln(xg, residual=r)
    y.float().sum().backward()          # torch's path: differentiable
    assert xg.grad is not None


def test_fused_layernorm_module_fallbacks():
    """Inputs the HIP kernel does not take run torch's own add and
    LayerNorm, bit for bit: x not 8-byte aligned, a residual that needs a
    gradient (parameters frozen)."""
    torch.manual_seed(4)
    ln = spattention.FusedLayerNorm(512).cuda().half()
    for p in ln.parameters():
        p.requires_grad_(False)
    buf = torch.randn(1 + 8 * 512, device="cuda").half()
    x = buf[1:].view(8, 512)                       # 2-byte offset
    assert x.data_ptr() % 8 != 0
    with torch.no_grad():
        assert torch.equal(ln(x), F.layer_norm(x, (512,), ln.weight, ln.bias, ln.eps))
    xa = torch.randn(8, 512, device="cuda").half()
    r = torch.randn(8, 512, device="cuda").half().requires_grad_(True)
    y = ln(xa, residual=r)
    assert y.requires_grad
    y.float().sum().backward()
    assert r.grad is not None
    with torch.no_grad():
        assert torch.equal(y, F.layer_norm(xa + r, (512,), ln.weight, ln.bias, ln.eps))
