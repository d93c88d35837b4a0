"""Programs written against the Olive C++ API run on the GPU and agree with
the oracle: the examples (olive_bfs, olive_sssp, olive_pagerank) and the
reference's own callers BFS.cu, SSSP.cu and PageRank.cu compiled unchanged
(ref_BFS, ref_SSSP, ref_PageRank).  The programs are built beforehand into
olive_amd/_apps/ (olive_amd.build.build_apps); a missing program is a skip
for the reference callers and a failure for the examples."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import DATA_FILES, ROOT, data_file

import oracle

APPS = os.path.join(ROOT, "olive_amd", "_apps")
INF = 0x7FFFFFFF
pytestmark = pytest.mark.gpu


def _run(name, args, cwd, required=True):
    exe = os.path.join(APPS, name)
    if not os.path.exists(exe):
        if required:
            pytest.fail(f"{exe} not built (olive_amd.build.build_apps)")
        pytest.skip(f"{name} not built (reference callers need the reference checkout at build time)")
    print(f"run {name} {' '.join(args)}", flush=True)
    r = subprocess.run([exe] + list(args), cwd=str(cwd), capture_output=True, text=True, timeout=90)
    return r


def _read_ints(path):
    return np.loadtxt(path, dtype=np.int64, ndmin=1)


def _csr(name):
    n, rp, col, w = oracle.load_edge_list(data_file(name))
    return n, rp, col, w


@pytest.mark.parametrize("name", DATA_FILES)
def test_example_bfs(name, tmp_path):
    n, rp, col, _ = _csr(name)
    if n == 0:
        pytest.skip("empty graph")
    r = _run("olive_bfs", [data_file(name), "-s", "0", "-o", "out.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    got = _read_ints(tmp_path / "out.txt")
    ref = oracle.bfs(rp, col, 0).astype(np.int64)
    ref[ref == INF] = -1
    np.testing.assert_array_equal(got, ref)


def _helpers_expected(rp, col, s):
    """Per BFS level: frontier size and the number of distinct out-neighbours
    of the frontier (the touched set), then reached / level sum / max level."""
    lv = oracle.bfs(rp, col, s).astype(np.int64)
    reached = lv != INF
    lines = []
    for L in range(int(lv[reached].max()) + 1):
        f = np.flatnonzero(lv == L)
        nb = np.concatenate([col[rp[u]:rp[u + 1]] for u in f]) if f.size else np.zeros(0, np.uint32)
        lines.append(f"level {L} frontier {f.size} touched {np.unique(nb).size}")
    lines.append(f"reached {int(reached.sum())} levelsum {int(lv[reached].sum())} maxlevel {int(lv[reached].max())}")
    return lines


@pytest.mark.parametrize("name", ["gridGraph_15", "starGraph_1K", "disorderedGraph_9", "maxflowGraph_100"])
def test_example_helpers(name, tmp_path):
    """getWorkqueueSize (dense and sparse), allVerticesInactive, the device
    vertexReduce and vertexTransform, checked against the oracle's BFS."""
    n, rp, col, _ = _csr(name)
    r = _run("olive_helpers", [data_file(name), "-s", "0"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines() == _helpers_expected(rp, col, 0)


@pytest.mark.parametrize("name", DATA_FILES)
def test_example_sssp_weighted(name, tmp_path):
    n, rp, col, w = _csr(name)
    if n == 0:
        pytest.skip("empty graph")
    r = _run("olive_sssp", [data_file(name), "-s", "0", "-o", "out.txt"], tmp_path)
    weights = w if w is not None else np.ones(col.size, dtype=np.int32)
    ref, rounds = oracle.sssp(rp, col, weights, 0)
    if rounds == -1:
        assert r.returncode == 4, r.stderr
        return
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(_read_ints(tmp_path / "out.txt"), ref.astype(np.int64))


@pytest.mark.parametrize("name", DATA_FILES)
def test_example_pagerank(name, tmp_path):
    n, rp, col, _ = _csr(name)
    if n == 0:
        pytest.skip("empty graph")
    r = _run("olive_pagerank", [data_file(name), "-max", "20", "-o", "out.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "out.txt", dtype=np.float64, ndmin=1)
    ref, _, _ = oracle.pagerank(rp, col, 20)
    # fp64 on both sides; only the order of the atomic sums differs.
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)


def test_example_bfs_rmat(tmp_path):
    """A larger graph through the examples: RMAT scale 14 written as an
    edge list, so that all three degree buckets are exercised."""
    n, rp, col = oracle.gen_rmat(14, 16, seed=3)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    path = tmp_path / "rmat14.txt"
    with open(path, "w") as fh:
        fh.write(f"{n} {col.size}\n")
        np.savetxt(fh, np.stack([src, col.astype(np.int64)], axis=1), fmt="%d")
    s = int(np.argmax(np.diff(rp)))
    r = _run("olive_bfs", [str(path), "-s", str(s), "-o", "out.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    ref = oracle.bfs(rp, col, s).astype(np.int64)
    ref[ref == INF] = -1
    np.testing.assert_array_equal(_read_ints(tmp_path / "out.txt"), ref)
    r = _run("olive_helpers", [str(path), "-s", str(s)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines() == _helpers_expected(rp, col, s)
    r = _run("olive_pagerank", [str(path), "-max", "20", "-o", "pr.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    ref_pr, _, _ = oracle.pagerank(rp, col, 20)
    np.testing.assert_allclose(np.loadtxt(tmp_path / "pr.txt"), ref_pr, rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("name", DATA_FILES)
def test_reference_bfs_caller(name, tmp_path):
    """BFS.cu unchanged; -round large enough for the 1000-vertex chain (its
    default of 100 truncates, BFS.cu:86)."""
    n, rp, col, _ = _csr(name)
    r = _run("ref_BFS", [data_file(name), "-round", "100000"], tmp_path, required=False)
    assert r.returncode == 0, r.stderr
    got = _read_ints(tmp_path / "BFS.txt") if n else np.zeros(0, np.int64)
    ref = oracle.bfs(rp, col, 0).astype(np.int64) if n else np.zeros(0, np.int64)
    np.testing.assert_array_equal(got[:n], ref)


@pytest.mark.parametrize("name", DATA_FILES)
def test_reference_sssp_caller(name, tmp_path):
    """SSSP.cu unchanged: it sets every weight to 1 (SSSP.cu:131), so its
    distances are the BFS levels."""
    n, rp, col, _ = _csr(name)
    if n == 0:
        pytest.skip("empty graph")
    r = _run("ref_SSSP", [data_file(name), "-round", "100000"], tmp_path, required=False)
    assert r.returncode == 0, r.stderr
    ref = oracle.bfs(rp, col, 0).astype(np.int64)
    np.testing.assert_array_equal(_read_ints(tmp_path / "SSSP.txt"), ref)


@pytest.mark.parametrize("name", DATA_FILES)
def test_reference_pagerank_caller(name, tmp_path):
    """PageRank.cu unchanged, `-round 19` = 20 updates (it runs N+1,
    PageRank.cu:122-133); it prints ranks with %f."""
    n, rp, col, _ = _csr(name)
    if n == 0:
        pytest.skip("empty graph")
    r = _run("ref_PageRank", [data_file(name), "-round", "19"], tmp_path, required=False)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "PageRank.txt", dtype=np.float64, ndmin=1)
    ref, _, _ = oracle.pagerank(rp, col, 20)
    np.testing.assert_allclose(got, ref, rtol=0, atol=5.1e-7)
