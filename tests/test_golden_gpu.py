"""The GPU operators on the golden fixtures (reference generator data,
unsigned lanes as the reference's MMAs compute them), prepared through the
library's own format code: outputs equal the fixtures bit for bit."""
import os

import numpy as np
import pytest
import torch

from magicube_amd import formats, sddmm, spmm

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "refdata.npz")


def test_spmm_on_golden_fixtures():
    d = np.load(GOLDEN, allow_pickle=False)
    m_vec, K, N = (int(x) for x in d["dims"])
    keys = sorted({k.rsplit("_", 1)[0] for k in d.files if k.startswith("spmm_") and k.endswith("_out")})
    assert len(keys) == 24
    for key in keys:
        pa, pb = (int(x) for x in key.split("_")[1][1:].split("R"))
        V = int(key.split("_V")[1])
        print("golden", key, flush=True)
        csr = formats.VectorCSR(m_vec, K, V, d[key + "_row_offsets"], d[key + "_col_indices"])
        dev = spmm.DeviceSpmm.upload(formats.prepare_spmm(csr, d[key + "_values"], pa, pb))
        rhs = torch.from_numpy(d[key + "_rhs"].view(np.int32).copy()).cuda()
        out = dev.matmul(rhs, N, sign=spmm.UNSIGNED)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), d[key + "_out"]), key


def test_sddmm_on_golden_fixtures():
    d = np.load(GOLDEN, allow_pickle=False)
    m_vec, K, N = (int(x) for x in d["dims"])
    keys = sorted({k.rsplit("_", 1)[0] for k in d.files if k.startswith("sddmm_") and k.endswith("_out")})
    assert len(keys) == 9
    for key in keys:
        bits = int(key.split("_")[1][:-1])
        V = int(key.split("_V")[1])
        print("golden", key, flush=True)
        csr = formats.VectorCSR(m_vec, N, V, d[key + "_row_offsets"], d[key + "_col_indices"])
        mask = sddmm.DeviceMask.upload(csr)
        lhs = torch.from_numpy(d[key + "_lhs"].view(np.int32).copy()).cuda()
        rhs = torch.from_numpy(d[key + "_rhs"].view(np.int32).copy()).cuda()
        out = mask.sddmm(bits, lhs, rhs, K, sign=spmm.UNSIGNED)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), d[key + "_out"]), key
