"""GPU parity for the device CSR ingest: COO -> CSR (stable radix sort by
source), the device transpose, and edge-list / DIMACS text parsed on the
device -- each equal, array for array, to the oracle's restated loaders
(oracle.load_edge_list / load_dimacs, csrGraph.h:106-236), numpy's stable
argsort, and the host transpose of olv_graph_create."""
import sys

import numpy as np
import pytest

from conftest import DATA_FILES, data_file

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected_csr(n, src, dst, w=None):
    """CSR of an edge list by numpy's stable sort on the source."""
    src = np.asarray(src, np.int64)
    order = np.argsort(src, kind="stable")
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, src + 1, 1)
    return (np.cumsum(rp).astype(np.uint32), np.asarray(dst, np.uint32)[order],
            None if w is None else np.asarray(w)[order])


def check_graph(olv, g, rp, col, w=None):
    h = g.to_host_csr()
    assert np.array_equal(h.row_ptr, rp)
    assert np.array_equal(h.col, col)
    if w is None:
        assert h.weights is None
    else:
        assert h.weights is not None and h.weights.dtype == w.dtype
        assert np.array_equal(h.weights.view(np.uint32), w.view(np.uint32))
    # the device transpose equals the host transpose of the same CSR
    ref = olv.Graph(olv.CsrGraph(rp, col))
    for a, b in zip(g.transpose_to_host(), ref.transpose_to_host()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,m,wt", [(1000, 20000, "i32"), (1000, 4096 * 3 + 17, "f32"), (70000, 50000, None),
                                    (1, 9, None), (7, 0, None), ((1 << 24) + 3, 5000, "i32")])
def test_coo_random(olv, n, m, wt):
    say(f"coo n={n} m={m} w={wt}")
    rng = np.random.default_rng(n + m)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = None
    if wt == "i32":
        w = rng.integers(-2 ** 31, 2 ** 31, m).astype(np.int32)
    elif wt == "f32":
        w = rng.standard_normal(m).astype(np.float32)
    g = olv.ingest_coo(n, src, dst, w)
    rp, col, we = expected_csr(n, src, dst, w)
    check_graph(olv, g, rp, col, we)


def test_coo_skewed_and_sorted(olv):
    """One hub source, long runs of empty rows, already-sorted input."""
    say("coo skewed")
    n = 100000
    src = np.concatenate([np.full(30000, n - 1), np.full(10, 3), np.arange(0, n, 97)]).astype(np.uint32)
    dst = (np.arange(src.size) * 7919 % n).astype(np.uint32)
    for s, d in [(src, dst), (np.sort(src, kind="stable"), dst)]:
        g = olv.ingest_coo(n, s, d)
        rp, col, _ = expected_csr(n, s, d)
        check_graph(olv, g, rp, col)


def test_coo_rmat20_shuffled(olv):
    """RMAT-20 edges in a random order: stable sort recovers each source's
    edges in input order."""
    say("coo rmat20 shuffled")
    csr = olv.gen_rmat(20, 16, seed=3)
    src = np.repeat(np.arange(csr.n, dtype=np.uint32), np.diff(csr.row_ptr.astype(np.int64)))
    perm = np.random.default_rng(3).permutation(csr.m)
    s, d = src[perm], csr.col[perm]
    g = olv.ingest_coo(csr.n, s, d)
    rp, col, _ = expected_csr(csr.n, s, d)
    assert np.array_equal(rp, csr.row_ptr)
    h = g.to_host_csr()
    assert np.array_equal(h.row_ptr, rp) and np.array_equal(h.col, col)
    irp, icol, ieid = g.transpose_to_host()
    # in-lists hold sources in ascending order; in_eid points back to the edge
    assert np.array_equal(irp[-1:], [csr.m])
    assert np.array_equal(h.col[ieid], np.repeat(np.arange(csr.n, dtype=np.uint32), np.diff(irp.astype(np.int64))))
    srcs = np.repeat(np.arange(csr.n, dtype=np.uint32), np.diff(h.row_ptr.astype(np.int64)))
    assert np.array_equal(srcs[ieid], icol)
    # within one in-list, repeated sources (duplicate edges) keep edge order
    lid = np.repeat(np.arange(csr.n), np.diff(irp.astype(np.int64)))
    same = (np.diff(icol.astype(np.int64)) == 0) & (np.diff(lid) == 0)
    assert np.all(np.diff(icol.astype(np.int64))[np.diff(lid) == 0] >= 0)
    assert np.all(np.diff(ieid.astype(np.int64))[same] > 0)


def test_coo_rejects_bad_ids(olv):
    with pytest.raises(olv.OliveError) as e:
        olv.ingest_coo(10, [1, 2, 10], [0, 0, 0])
    assert e.value.code == 1


@pytest.mark.parametrize("name", DATA_FILES)
def test_edge_list_files(olv, name):
    """The reference's data files, read from disk and parsed on the device."""
    say(f"edge list file {name}")
    n, rp, col, w = oracle.load_edge_list(data_file(name))
    g = olv.ingest_file(data_file(name))
    check_graph(olv, g, rp, col, w)
    with open(data_file(name), "rb") as fh:
        g2 = olv.ingest_edge_list_text(fh.read())
    check_graph(olv, g2, rp, col, w)


def _messy(text: bytes) -> bytes:
    """Comments, blank lines, CRLF, tabs and leading blanks around the same
    records."""
    out = [b"# a comment line\n", b"\n", b"   # indented comment\n"]
    for i, ln in enumerate(text.split(b"\n")):
        if not ln:
            continue
        if i % 5 == 1:
            ln = b"  " + ln.replace(b" ", b"\t")
        if i % 7 == 2:
            ln = ln + b"\r"
        out.append(ln + b"\n")
        if i % 11 == 3:
            out.append(b"#\n\n")
    return b"".join(out)


@pytest.mark.parametrize("weights", [False, True])
def test_edge_list_generated(olv, tmp_path, weights):
    say(f"edge list generated weights={weights}")
    csr = olv.gen_rmat(14, 8, seed=5)
    src = np.repeat(np.arange(csr.n), np.diff(csr.row_ptr.astype(np.int64)))
    perm = np.random.default_rng(5).permutation(csr.m)
    w = np.random.default_rng(6).integers(-1000, 1000, csr.m) if weights else None
    text = _messy(olv.format_edge_list(csr.n, src[perm], csr.col[perm], None if w is None else w[perm]))
    p = tmp_path / "g.txt"
    p.write_bytes(text)
    n, rp, col, wo = oracle.load_edge_list(str(p))
    assert (wo is not None) == weights
    check_graph(olv, olv.ingest_edge_list_text(text), rp, col, wo)


def test_edge_list_errors(olv):
    bad = {
        b"3 2\n0 1\n1 x\n": "not an integer",
        b"3 2\n0 1\n1 3\n": "out of range",
        b"3 2\n0 1\n1\n": "fewer than two",
        b"3 2\n0 1\n": "edge lines",
        b"# only comments\n": "header",
    }
    for text, msg in bad.items():
        with pytest.raises(olv.OliveError) as e:
            olv.ingest_edge_list_text(text)
        assert e.value.code == 1 and msg in str(e.value), (text, str(e.value))
    # extra columns after the third are ignored; a partial third column drops weights
    g = olv.ingest_edge_list_text(b"3 2\n0 1 5 junk\n1 2 6\n")
    assert np.array_equal(g.to_host_csr().weights, [5, 6])
    g = olv.ingest_edge_list_text(b"3 2\n0 1 5\n1 2\n")
    assert g.to_host_csr().weights is None


def _dimacs(csr, comments=True) -> bytes:
    lines = [f"{csr.n} {csr.m // 2}\n".encode()]
    if comments:
        lines.insert(0, b"# generated\n")
    for v in range(csr.n):
        nb = csr.col[csr.row_ptr[v]:csr.row_ptr[v + 1]] + 1
        lines.append((" ".join(str(x) for x in nb) + "\n").encode())
        if comments and v % 13 == 4:
            lines.append(b"# between vertices\n")
    return b"".join(lines)


def test_dimacs_generated(olv, tmp_path):
    say("dimacs generated")
    csr = olv.gen_grid(37, 41)  # symmetric: 2m arcs
    text = _dimacs(csr)
    p = tmp_path / "g.graph"
    p.write_bytes(text)
    n, rp, col = oracle.load_dimacs(str(p))
    check_graph(olv, olv.ingest_dimacs_text(text), rp, col)
    check_graph(olv, olv.ingest_file(str(p), fmt="dimacs"), rp, col)
    # trailing vertices without lines are empty rows
    short = b"4 1\n2\n1\n"
    p2 = tmp_path / "s.graph"
    p2.write_bytes(short)
    n, rp, col = oracle.load_dimacs(str(p2))
    check_graph(olv, olv.ingest_dimacs_text(short), rp, col)


def test_dimacs_errors(olv):
    for text in [b"3 2\n2\n1\n", b"2 1\n2\n5\n", b"2 1\n2\nx\n"]:
        with pytest.raises(olv.OliveError) as e:
            olv.ingest_dimacs_text(text)
        assert e.value.code == 1


def test_ingested_graph_runs_bfs(olv):
    """A graph ingested on the device serves the algorithms directly."""
    csr = olv.gen_rmat(16, 16, seed=9)
    src = np.repeat(np.arange(csr.n), np.diff(csr.row_ptr.astype(np.int64)))
    g = olv.ingest_edge_list_text(olv.format_edge_list(csr.n, src, csr.col))
    s = int(np.flatnonzero(csr.outdeg())[0])
    got = g.bfs(s)[0].numpy()
    assert np.array_equal(got, oracle.bfs(csr.row_ptr, csr.col, s))
