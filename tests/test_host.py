"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, the seeded generators match their specification and
are deterministic, and the product's file loaders match the oracle's."""
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import DATA_FILES, ROOT, data_file

import olive_amd
from olive_amd import _lib


def header_symbols():
    hdr = open(os.path.join(ROOT, "include", "olive_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(olv_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)


def test_library_has_gfx950_code_object():
    so = _lib.LIB_PATH
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o", f"--input={so}"],
                         capture_output=True, text=True)
    blob = open(so, "rb").read()
    assert b"gfx950" in blob, "no gfx950 code object in libolive_hip.so"


def test_version_and_last_error():
    L = _lib.lib()
    assert L.olv_version() >= 100
    # a failing host-side call leaves a message (no device needed)
    assert L.olv_gen_grid(0, 5, None) != 0
    assert b"" != L.olv_last_error()


def csr_sha(g):
    h = hashlib.sha256()
    h.update(g.row_ptr.tobytes())
    h.update(g.col.tobytes())
    if g.weights is not None:
        h.update(g.weights.tobytes())
    return h.hexdigest()[:16]


def test_grid_matches_spec_and_reference_layout():
    g = olive_amd.gen_grid(3, 5)
    n, rp, col, _ = oracle.load_edge_list(data_file("gridGraph_15"))
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)
    g = olive_amd.gen_grid(128, 128)
    n, rp, col = oracle.gen_grid(128, 128)
    assert g.n == 16384 and g.m == 65024
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)


@pytest.mark.parametrize("scale,ef,seed", [(8, 16, 1), (10, 16, 7), (12, 8, 3)])
def test_rmat_matches_spec(scale, ef, seed):
    g = olive_amd.gen_rmat(scale, ef, seed=seed)
    n, rp, col = oracle.gen_rmat(scale, ef, seed=seed)
    assert g.n == n
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)


def test_rmat_structure():
    g = olive_amd.gen_rmat(16, 16, seed=1)
    n = g.n
    src = np.repeat(np.arange(n), g.outdeg())
    assert np.all(src != g.col), "self-loops must be dropped"
    assert g.col.max() < n
    m_gen = n * 16
    # expected self-loops: m_gen * (a+d)^scale
    exp_loops = m_gen * 0.62 ** 16
    assert abs((m_gen - g.m) - exp_loops) < 6 * np.sqrt(exp_loops) + 2
    top = n // 2
    assert abs((src < top).mean() - 0.76) < 0.01
    assert abs((g.col < top).mean() - 0.76) < 0.01
    # rows sorted by destination
    for v in range(0, n, 997):
        r = g.col[g.row_ptr[v]:g.row_ptr[v + 1]]
        assert np.all(r[1:] >= r[:-1])


def test_uniform_and_weights_match_spec():
    g = olive_amd.gen_uniform(3000, 20000, seed=5)
    n, rp, col = oracle.gen_uniform(3000, 20000, seed=5)
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)
    gw = olive_amd.with_weights_i32(g, 1, 64, seed=5)
    assert np.array_equal(gw.weights, oracle.weights_i32(g.m, 1, 64, 5))
    assert gw.weights.min() >= 1 and gw.weights.max() <= 64
    gf = olive_amd.with_weights_f32_normalized(olive_amd.gen_rmat(8, 8, seed=2), seed=9)
    ref = oracle.weights_f32_normalized(gf.row_ptr, 9)
    assert np.array_equal(gf.weights, ref)
    sums = np.add.reduceat(gf.weights.astype(np.float64), gf.row_ptr[:-1][np.diff(gf.row_ptr) > 0])
    assert np.allclose(sums, 1.0, atol=1e-6)
    t = olive_amd.gen_dirichlet(5000, seed=3)
    assert t.dtype == np.float32 and np.all(t > 0) and abs(t.astype(np.float64).sum() - 1) < 1e-5


def test_generators_deterministic_across_thread_counts():
    code = ("import olive_amd,hashlib;g=olive_amd.gen_rmat(14,16,seed=3);"
            "u=olive_amd.gen_uniform(50000,400000,seed=4);"
            "print(hashlib.sha256(g.row_ptr.tobytes()+g.col.tobytes()+u.row_ptr.tobytes()+u.col.tobytes()).hexdigest())")
    outs = []
    for t in ("1", "3", "8"):
        env = dict(os.environ, OLV_THREADS=t, PYTHONPATH=ROOT)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                   check=True).stdout.strip())
    assert len(set(outs)) == 1, outs


def test_generator_fingerprints():
    """Committed SHA-256 prefixes of small generator outputs: any change to
    the generators' output for a seed shows up here."""
    fp = os.path.join(ROOT, "tests", "golden", "generator_sha.txt")
    cur = {
        "rmat10_ef16_s7": csr_sha(olive_amd.gen_rmat(10, 16, seed=7)),
        "rmat14_ef16_s1": csr_sha(olive_amd.gen_rmat(14, 16, seed=1)),
        "grid128": csr_sha(olive_amd.gen_grid(128, 128)),
        "unif3000_s5_w": csr_sha(olive_amd.with_weights_i32(olive_amd.gen_uniform(3000, 20000, seed=5), 1, 64, 5)),
    }
    want = dict(line.split() for line in open(fp) if line.strip())
    assert cur == want


def test_batch_scales_in_range():
    sc = [olive_amd.batch_scale(1, i, 12, 18) for i in range(64)]
    assert min(sc) >= 12 and max(sc) <= 18 and len(set(sc)) > 3
    assert len({olive_amd.batch_seed(1, i) for i in range(64)}) == 64


@pytest.mark.parametrize("name", DATA_FILES)
def test_product_loader_matches_oracle(name):
    g = olive_amd.load_edge_list(data_file(name))
    n, rp, col, w = oracle.load_edge_list(data_file(name))
    assert g.n == n and np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)
    if w is None:
        assert g.weights is None
    else:
        assert np.array_equal(g.weights, w)


def test_product_loader_dimacs(tmp_path):
    p = tmp_path / "g.metis"
    p.write_text("# comment\n4 3\n2 3\n1\n1 4\n3\n")
    g = olive_amd.load_dimacs(str(p))
    n, rp, col = oracle.load_dimacs(str(p))
    assert g.n == 4 and np.array_equal(g.row_ptr, rp) and np.array_equal(g.col, col)
    assert g.col.tolist() == [1, 2, 0, 0, 3, 2]


def test_loader_rejects_bad_files(tmp_path):
    p = tmp_path / "bad"
    p.write_text("3 2\n0 1\n")
    with pytest.raises(olive_amd.OliveError):
        olive_amd.load_edge_list(str(p))
    p.write_text("3 1\n0 7\n")
    with pytest.raises(olive_amd.OliveError):
        olive_amd.load_edge_list(str(p))


def test_autograd_wrapper_validates_inputs():
    """The torch wrapper rejects host tensors, other dtypes and wrong shapes
    before any library call (no GPU needed)."""
    torch = pytest.importorskip("torch")
    from olive_amd import autograd

    with pytest.raises(TypeError):
        autograd._check_vec(torch.zeros(4), 4, "w")          # host tensor
    with pytest.raises(TypeError):
        autograd._check_vec(np.zeros(4, np.float32), 4, "w")  # not a tensor
    assert autograd.PageRankFunction._fn is None or callable(autograd.PageRankFunction._fn.apply)


def test_gpus_flag_other_than_one_exits(tmp_path):
    """bench.py runs one GPU; any other --gpus value exits with a message."""
    import json
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_format_edge_list_round_trip(tmp_path):
    """format_edge_list writes the reference's text format: the oracle's
    loader reads back the same edges (signed weights included)."""
    rng = np.random.default_rng(11)
    n, m = 5000, 20000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(-2 ** 31, 2 ** 31, m)
    p = tmp_path / "e.txt"
    p.write_bytes(olive_amd.format_edge_list(n, src, dst, w, comment="round trip"))
    nn, rp, col, ww = oracle.load_edge_list(str(p))
    order = np.argsort(src, kind="stable")
    assert nn == n and rp[-1] == m
    assert np.array_equal(col, dst[order]) and np.array_equal(ww, w[order].astype(np.int32))
    assert olive_amd.format_edge_list(3, [], []) == b"3 0\n"
