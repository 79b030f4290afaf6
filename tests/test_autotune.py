"""Autotuner host logic: cache keys and the on-disk cache (no GPU)."""
import json

from magicube_amd import autotune, formats, synthetic


class _Dev:
    def __init__(self, prep):
        self.prep = prep


def _prep(m=256, k=512, V=8, s=0.9, bits=(8, 8)):
    csr = synthetic.uniform_block_mask(m, k, V, s, seed=1)
    a = synthetic.random_lanes(csr.nnz_vec * V, bits[0], True, 2)
    return formats.prepare_spmm(csr, formats.pack_lanes(a, bits[0]), *bits)


def test_key_separates_shapes_and_buckets_sizes():
    p = _prep()
    k1 = autotune.spmm_key(_Dev(p), 512, 0, "gfx950")
    assert k1.startswith("spmm|gfx950|L8R8|s|V8|N512|")
    assert autotune.spmm_key(_Dev(p), 256, 0, "gfx950") != k1
    assert autotune.spmm_key(_Dev(p), 512, 1, "gfx950") != k1
    assert autotune.spmm_key(_Dev(_prep(bits=(8, 4))), 512, 0, "gfx950") != k1
    # same size buckets for a slightly different pattern of the same shape
    assert autotune.spmm_key(_Dev(_prep(s=0.91)), 512, 0, "gfx950") == k1


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("MC_TUNE_CACHE", str(path))
    assert autotune.load_cache() == {}
    autotune.save_cache({"k": {"config": 3, "us": 1.5}})
    assert json.loads(path.read_text()) == {"k": {"config": 3, "us": 1.5}}
    assert autotune.load_cache() == {"k": {"config": 3, "us": 1.5}}
    path.write_text("not json")
    assert autotune.load_cache() == {}
