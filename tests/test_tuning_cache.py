"""The tuning cache and graph signatures (host code only, no GPU): a stored
olv_tuning reads back field for field, a key's line is replaced in place,
invalid settings are refused, and the signature tells graphs apart by size
and degree histogram while equal graphs share it."""
import numpy as np
import pytest

import olive_amd as olv
from olive_amd import _lib


def _tuning(**kw):
    t = _lib.Tuning()
    base = dict(bfs_expand_per_cu=8, bfs_expand_per_cu_twin=6, bfs_compact_per_cu=8, td_wave_thr=64,
                td_block_thr=256, hub_thr=1024, alpha=14.0, beta=24.0, mbfs_pull_per_cu=48, seg_grid_per_cu=32,
                seg_unroll=8, sssp_relax_per_cu=6, sssp_delta_num=32)
    bounds = kw.pop("seg_bounds", [8, 96, 1024])
    base.update(kw)
    for k, v in base.items():
        setattr(t, k, v)
    t.seg_bounds[:] = bounds
    return t


def test_cache_round_trip_and_replace(tmp_path):
    path = tmp_path / "tune.jsonl"
    assert olv.tuning_cache_load(path, "v1-e1-h0") is None  # no file yet
    a = _tuning(bfs_expand_per_cu_twin=12, alpha=6.5, hub_thr=2048, seg_bounds=[4, 48, 2048], seg_unroll=16)
    b = _tuning(mbfs_pull_per_cu=96, beta=48.0)
    olv.tuning_cache_store(path, "key-a", a)
    olv.tuning_cache_store(path, "key-b", b)
    ra, rb = olv.tuning_cache_load(path, "key-a"), olv.tuning_cache_load(path, "key-b")
    for r, x in ((ra, a), (rb, b)):
        for name, _ in _lib.Tuning._fields_:
            if name in ("source", "tune_ms"):
                continue
            if name == "seg_bounds":
                assert list(r.seg_bounds) == list(x.seg_bounds)
            else:
                assert getattr(r, name) == pytest.approx(getattr(x, name)), name
        assert r.source == 2
    a2 = _tuning(bfs_compact_per_cu=16)
    olv.tuning_cache_store(path, "key-a", a2)
    lines = path.read_text().splitlines()
    assert len(lines) == 2 and sum('"key": "key-a"' in l for l in lines) == 1
    assert olv.tuning_cache_load(path, "key-a").bfs_compact_per_cu == 16
    assert olv.tuning_cache_load(path, "key-c") is None


@pytest.mark.parametrize("field,value", [("seg_unroll", 5), ("hub_thr", 100), ("td_block_thr", 32),
                                         ("bfs_expand_per_cu", 0), ("alpha", 0.0)])
def test_cache_refuses_invalid(tmp_path, field, value):
    t = _tuning(**{field: value})
    with pytest.raises(olv.OliveError) as e:
        olv.tuning_cache_store(tmp_path / "t.jsonl", "k", t)
    assert e.value.code == 1


def test_cache_malformed_line(tmp_path):
    path = tmp_path / "t.jsonl"
    path.write_text('{"key": "k", "bfs_expand_per_cu": 8}\n')
    with pytest.raises(olv.OliveError):
        olv.tuning_cache_load(path, "k")


def test_signature():
    a = olv.gen_rmat(12, 16, seed=1)
    b = olv.gen_rmat(12, 16, seed=1)
    c = olv.gen_rmat(12, 16, seed=2)
    g = olv.gen_grid(64, 64)
    sa, sb, sc, sg = (olv.csr_signature(x) for x in (a, b, c, g))
    assert sa == sb
    assert sa.startswith(f"v{a.n}-e{a.m}-h")
    assert sg != sa and sc != sa
    # the same degree histograms and counts give the same key (relabelled copy)
    n = g.n
    perm = np.random.default_rng(0).permutation(n)
    src = perm[np.repeat(np.arange(n), g.outdeg())]
    dst = perm[g.col]
    o = np.argsort(src, kind="stable")
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(src, minlength=n))
    assert olv.csr_signature(olv.CsrGraph(rp, dst[o])) == sg
