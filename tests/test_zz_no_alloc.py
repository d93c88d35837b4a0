"""olv_bfs_noalloc -- BFS whose work buffers live in a static device array
(no allocation in the call) -- against olv_bfs in the graph's own numbering
and the oracle, on two calls in a row (the second on a smaller graph), and
its argument checks."""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(olv, g, src, out, flags=0):
    from olive_amd import _lib
    st = _lib.Stats()
    _lib.check(_lib.lib().olv_bfs_noalloc(g.h, src, out.c, flags, ctypes.byref(st)), "olv_bfs_noalloc")
    return st


def test_noalloc_matches_bfs_two_calls(olv):
    from olive_amd import _lib
    cases = [(olv.gen_rmat(18, 16, seed=3), 0), (olv.gen_rmat(12, 16, seed=8), 0),
             (olv.gen_rmat(12, 16, seed=8), _lib.OLV_NO_DIRECTION_OPT), (olv.gen_grid(40, 50), _lib.OLV_HOST_LOOP),
             (olv.gen_rmat(14, 16, seed=2), _lib.OLV_FORCE_BOTTOM_UP)]
    for csr, flags in cases:
        say(f"noalloc n={csr.n} m={csr.m} flags={flags:#x}")
        g = olv.Graph(csr)
        nz = np.flatnonzero(csr.outdeg())
        for s in (int(nz[0]), int(np.argmax(csr.outdeg())), int(nz[-1])):
            ref = oracle.bfs(csr.row_ptr, csr.col, s)
            plain, pst = g.bfs(s, flags=flags | _lib.OLV_GRAPH_ORDER)
            out = olv.DeviceBuffer(csr.n, np.int32)
            out.fill(0x5A5A5A5A)
            st = _noalloc(olv, g, s, out, flags)
            got = out.numpy()
            assert np.array_equal(got, ref), (csr.n, s, flags)
            assert np.array_equal(got, plain.numpy())
            assert st.iterations == pst.iterations and st.vertices_reached == pst.vertices_reached
            assert st.edges_traversed == pst.edges_traversed


def test_noalloc_rejects(olv):
    from olive_amd import _lib
    L = _lib.lib()
    csr = olv.gen_rmat(10, 8, seed=1)
    g = olv.Graph(csr)
    out = olv.DeviceBuffer(csr.n, np.int32)
    assert L.olv_bfs_noalloc(g.h, csr.n, out.c, 0, None) == 1  # source out of range
    assert L.olv_bfs_noalloc(g.h, 0, out.c, _lib.OLV_PTR_HOST, None) == 1
    assert L.olv_bfs_noalloc(g.h, 0, out.c, _lib.OLV_PERSISTENT, None) == 1
    g2 = olv.Graph(csr, transpose=False)  # no in-edge CSR: push-only runs
    assert L.olv_bfs_noalloc(g2.h, 0, out.c, 0, None) == 1
    assert b"OLV_BUILD_TRANSPOSE" in L.olv_last_error()
    st = _noalloc(olv, g2, 0, out, _lib.OLV_NO_DIRECTION_OPT)
    assert np.array_equal(out.numpy(), oracle.bfs(csr.row_ptr, csr.col, 0))
    # a graph whose work memory exceeds the static array: a clear error
    n = 10_000_000
    rp = np.ones(n + 1, np.uint32)
    rp[0] = 0
    big = olv.Graph(olv.CsrGraph(rp, np.array([1], np.uint32)), transpose=False)
    big_out = olv.DeviceBuffer(n, np.int32)
    assert L.olv_bfs_noalloc(big.h, 0, big_out.c, _lib.OLV_NO_DIRECTION_OPT, None) == 4
    assert b"static scratch" in L.olv_last_error()
    # the array still works afterwards
    st = _noalloc(olv, g, 5, out)
    assert np.array_equal(out.numpy(), oracle.bfs(csr.row_ptr, csr.col, 5))
