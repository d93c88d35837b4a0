"""olive_amd -- an MI355X (gfx950) frontier engine with Olive's graph API.

Python host layer over the C ABI of ``libolive_hip.so``
(``include/olive_hip.h``).  The C++ API (``Olive<V,E,A>`` / ``Oliver``) lives
in ``olive_amd/include``; this module serves tests, benchmarks and the torch
autograd wrapper.  Every compute call runs the HIP kernels; if the library is
missing the import of a compute function raises.

Naming follows the reference: graphs are CSR ``row_ptr`` (the reference's
``vertices``) and ``col`` (``edges``), csrGraph.h:61-63; BFS levels and SSSP
distances use INF = 0x7fffffff (BFS.cu:101, SSSP.cu:123).
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref

import numpy as np

from ._lib import (OLV_BUILD_TRANSPOSE, OLV_FMT_DIMACS, OLV_FMT_EDGE_LIST, OLV_FORCE_BOTTOM_UP, OLV_GRAPH_ORDER,
                   OLV_HOST_LOOP, OLV_NO_DIRECTION_OPT, OLV_PERSISTENT, OLV_PROFILE, OLV_PTR_HOST, OLV_W_F32, OLV_W_I32, OLV_W_NONE,
                   HostCsr, OliveError, Stats, Tuning, check, lib)

INF = 0x7FFFFFFF

__all__ = [
    "INF", "OliveError", "Stats", "CsrGraph", "Graph", "DeviceBuffer", "init", "gen_grid", "gen_rmat",
    "gen_uniform", "gen_dirichlet", "batch_scale", "batch_seed", "load_edge_list", "load_dimacs",
    "ingest_coo", "ingest_edge_list_text", "ingest_dimacs_text", "ingest_file", "format_edge_list",
    "write_text", "write_binary", "read_binary",
    "OLV_HOST_LOOP", "OLV_NO_DIRECTION_OPT", "OLV_FORCE_BOTTOM_UP", "OLV_PTR_HOST", "OLV_GRAPH_ORDER",
]

_init_done = False
# Live device objects, released by _release_all at interpreter exit while the
# HIP runtime is still up; after that (and during interpreter finalisation)
# no native free is attempted.
_live: "weakref.WeakSet" = weakref.WeakSet()
_released = False


def _release_all():
    """atexit hook: wait for queued work, then free every live graph and
    buffer in this order (graphs first), before Python's own teardown and the
    runtime's exit handlers run."""
    global _released
    if _released or not _init_done:
        _released = True
        return
    try:
        lib().olv_synchronize()
        objs = list(_live)
        for o in objs:
            if isinstance(o, Graph):
                o.close()
        for o in objs:
            if isinstance(o, DeviceBuffer):
                o.free()
    finally:
        _released = True


def _native_free_allowed():
    return not _released and not sys.is_finalizing()


def init(device: int = 0):
    """Select the HIP device (olv_init)."""
    global _init_done
    if not _init_done:
        check(lib().olv_init(device), "olv_init")
        _init_done = True
        atexit.register(_release_all)


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


# --------------------------------------------------------------- buffers
class DeviceBuffer:
    """Device allocation from olv_malloc holding a 1-D array of ``dtype``."""

    def __init__(self, count: int, dtype):
        init()
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().olv_malloc(ctypes.byref(p), max(self.nbytes, 4)), "olv_malloc")
        self.ptr = p.value
        _live.add(self)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.size, a.dtype)
        if a.size:
            check(lib().olv_memcpy_h2d(b.ptr, a.ctypes.data, b.nbytes), "olv_memcpy_h2d")
        return b

    def numpy(self):
        out = np.empty(self.count, dtype=self.dtype)
        if self.count:
            check(lib().olv_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "olv_memcpy_d2h")
        return out

    def fill(self, value: int):
        check(lib().olv_fill(self.ptr, int(value), self.count, self.dtype.itemsize), "olv_fill")

    @property
    def c(self):
        return ctypes.c_void_p(self.ptr)

    def free(self):
        if getattr(self, "ptr", None):
            if _native_free_allowed():
                lib().olv_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------------- graphs
class CsrGraph:
    """Host CSR (the reference's CsrGraph<int,int>, csrGraph.h:54-84) plus
    optional per-edge weights aligned with ``col``."""

    def __init__(self, row_ptr, col, weights=None):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.weights = None if weights is None else np.ascontiguousarray(weights)
        if self.weights is not None and self.weights.dtype not in (np.int32, np.float32):
            raise ValueError("weights must be int32 or float32")

    @property
    def n(self):
        return int(self.row_ptr.size - 1)

    @property
    def m(self):
        return int(self.col.size)

    def outdeg(self):
        return np.diff(self.row_ptr.astype(np.int64))

    def with_weights(self, w):
        return CsrGraph(self.row_ptr, self.col, w)


def csr_signature(g: CsrGraph) -> str:
    """olv_csr_signature of a host CSR (the tuning cache key)."""
    buf = ctypes.create_string_buffer(96)
    check(lib().olv_csr_signature(g.n, g.m, g.row_ptr.ctypes.data, _np_ptr(g.col), buf, 96), "olv_csr_signature")
    return buf.value.decode()


def tuning_cache_load(path, key: str):
    """The cached olv_tuning of `key`, or None."""
    t, found = Tuning(), ctypes.c_int()
    check(lib().olv_tuning_cache_load(str(path).encode(), key.encode(), ctypes.byref(t), ctypes.byref(found)),
          "olv_tuning_cache_load")
    return t if found.value else None


def tuning_cache_store(path, key: str, t: Tuning):
    check(lib().olv_tuning_cache_store(str(path).encode(), key.encode(), ctypes.byref(t)), "olv_tuning_cache_store")


def _take_host_csr(h: HostCsr) -> CsrGraph:
    n, m = h.n, h.m
    rp = np.ctypeslib.as_array(h.row_ptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(m, 1),))[:m].copy()
    w = None
    if h.weights and h.weight_dtype in (OLV_W_I32, OLV_W_F32):
        ct = ctypes.c_int32 if h.weight_dtype == OLV_W_I32 else ctypes.c_float
        arr = ctypes.cast(h.weights, ctypes.POINTER(ct))
        w = np.ctypeslib.as_array(arr, shape=(max(m, 1),))[:m].copy()
    lib().olv_host_csr_free(ctypes.byref(h))
    return CsrGraph(rp, col, w)


def gen_grid(rows: int, cols: int) -> CsrGraph:
    """Directed R x C grid, id = C*row + col, edges to the 4 neighbours."""
    h = HostCsr()
    check(lib().olv_gen_grid(rows, cols, ctypes.byref(h)), "olv_gen_grid")
    return _take_host_csr(h)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
             drop_self_loops: bool = True) -> CsrGraph:
    h = HostCsr()
    check(lib().olv_gen_rmat(scale, edge_factor, a, b, c, seed, int(drop_self_loops), ctypes.byref(h)),
          "olv_gen_rmat")
    return _take_host_csr(h)


def gen_uniform(n: int, m: int, seed: int = 1) -> CsrGraph:
    h = HostCsr()
    check(lib().olv_gen_uniform(n, m, seed, ctypes.byref(h)), "olv_gen_uniform")
    return _take_host_csr(h)


def _to_host_csr(g: CsrGraph) -> HostCsr:
    # borrowed view for the weight generators (arrays owned by numpy)
    h = HostCsr()
    h.n, h.m = g.n, g.m
    h.row_ptr = g.row_ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    h.col = g.col.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    h.weights = None
    h.weight_dtype = OLV_W_NONE
    return h


def _attach(g: CsrGraph, fn, *args) -> CsrGraph:
    h = _to_host_csr(g)
    check(fn(ctypes.byref(h), *args), fn.__name__)
    ct = ctypes.c_int32 if h.weight_dtype == OLV_W_I32 else ctypes.c_float
    arr = ctypes.cast(h.weights, ctypes.POINTER(ct))
    w = np.ctypeslib.as_array(arr, shape=(max(g.m, 1),))[:g.m].copy()
    # the library allocated the weights; free only them
    h.row_ptr = ctypes.POINTER(ctypes.c_uint32)()
    h.col = ctypes.POINTER(ctypes.c_uint32)()
    lib().olv_host_csr_free(ctypes.byref(h))
    return g.with_weights(w)


def with_weights_i32(g: CsrGraph, lo=1, hi=64, seed=1) -> CsrGraph:
    return _attach(g, lib().olv_gen_weights_i32, lo, hi, seed)


def with_weights_f32_normalized(g: CsrGraph, seed=1) -> CsrGraph:
    return _attach(g, lib().olv_gen_weights_f32_normalized, seed)


def gen_dirichlet(n: int, seed: int = 1):
    out = np.empty(n, dtype=np.float32)
    check(lib().olv_gen_dirichlet(n, seed, out.ctypes.data), "olv_gen_dirichlet")
    return out


def batch_scale(seed, i, lo, hi):
    return int(lib().olv_gen_batch_scale(seed, i, lo, hi))


def batch_seed(seed, i):
    return int(lib().olv_gen_batch_seed(seed, i))


# ---------------------------------------------------------- device ingest
def _ndigits(a):
    nd = np.ones(a.shape, np.int64)
    for k in range(1, 19):
        nd += a >= 10 ** k
    return nd


def format_edge_list(n: int, src, dst, weights=None, comment: str | None = None) -> bytes:
    """Edge-list text in the reference's format (csrGraph.h:106-171): an
    optional '#' comment line, the "n m" header, then one "src dst [w]" line
    per edge in the given order.  Vectorised, for inputs of many millions of
    edges."""
    cols = [np.asarray(src, np.int64), np.asarray(dst, np.int64)]
    if weights is not None:
        cols.append(np.asarray(weights, np.int64))
    m = cols[0].size
    head = ((f"# {comment}\n" if comment else "") + f"{n} {m}\n").encode()
    neg = [c < 0 for c in cols]
    mag = [np.abs(c) for c in cols]
    nds = [_ndigits(a) for a in mag]
    length = sum(nd + ng for nd, ng in zip(nds, neg)) + len(cols)  # separators and newline
    starts = np.empty(m, np.int64)
    if m:
        starts[0] = 0
        np.cumsum(length[:-1], out=starts[1:])
    total = int(starts[-1] + length[-1]) if m else 0
    out = np.empty(total, np.uint8)
    pos = starts
    for c, (ng, a, nd) in enumerate(zip(neg, mag, nds)):
        out[pos[ng]] = ord("-")
        pos = pos + ng
        for k in range(int(nd.max()) if m else 0):
            sel = k < nd
            out[pos[sel] + nd[sel] - 1 - k] = ord("0") + (a[sel] // 10 ** k) % 10
        pos = pos + nd
        out[pos] = ord(" ") if c + 1 < len(cols) else ord("\n")
        pos = pos + 1
    return head + out.tobytes()


def ingest_coo(n: int, src, dst, weights=None, transpose: bool = True) -> Graph:
    """Edge list in any order -> resident graph, built on the device
    (olv_graph_from_coo: stable sort by source, edges of a source keep their
    order; weights follow their edges)."""
    init()
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    if s.size != d.size:
        raise ValueError("src and dst differ in length")
    wd, wp, w = OLV_W_NONE, None, None
    if weights is not None:
        w = np.ascontiguousarray(weights)
        if w.dtype not in (np.int32, np.float32) or w.size != s.size:
            raise ValueError("weights must be int32 or float32, one per edge")
        wd = OLV_W_I32 if w.dtype == np.int32 else OLV_W_F32
        wp = w.ctypes.data
    h = ctypes.c_void_p()
    flags = OLV_PTR_HOST | (OLV_BUILD_TRANSPOSE if transpose else 0)
    check(lib().olv_graph_from_coo(ctypes.byref(h), n, s.size, s.ctypes.data if s.size else None,
                                   d.ctypes.data if d.size else None, wp, wd, flags, None), "olv_graph_from_coo")
    return Graph._adopt(h)


def _ingest_text(fn, name, text, transpose):
    init()
    buf = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    flags = OLV_PTR_HOST | (OLV_BUILD_TRANSPOSE if transpose else 0)
    check(fn(ctypes.byref(h), buf, len(buf), flags, None), name)
    return Graph._adopt(h)


def ingest_edge_list_text(text, transpose: bool = True) -> Graph:
    """Edge-list text (the format of load_edge_list) parsed on the device."""
    return _ingest_text(lib().olv_graph_from_edge_list_text, "olv_graph_from_edge_list_text", text, transpose)


def ingest_dimacs_text(text, transpose: bool = True) -> Graph:
    """DIMACS adjacency text (the format of load_dimacs) parsed on the device."""
    return _ingest_text(lib().olv_graph_from_dimacs_text, "olv_graph_from_dimacs_text", text, transpose)


def ingest_file(path, fmt: str = "edge_list", transpose: bool = True) -> Graph:
    """A file read into pinned memory and parsed on the device
    (olv_graph_load_file); fmt is "edge_list" or "dimacs"."""
    init()
    f = {"edge_list": OLV_FMT_EDGE_LIST, "dimacs": OLV_FMT_DIMACS}[fmt]
    h = ctypes.c_void_p()
    check(lib().olv_graph_load_file(ctypes.byref(h), str(path).encode(), f,
                                    OLV_BUILD_TRANSPOSE if transpose else 0, None), "olv_graph_load_file")
    return Graph._adopt(h)


# ----------------------------------------------------------- result egress
_TXT = {"%d": 0, "%f32": 1, "%f": 2}


def write_text(path, values, fmt: str = "%d", stride: int = 0, stats: Stats | None = None, profile: bool = False):
    """One value per line in the reference programs' formats ("%d" for int32,
    "%f" for float64, "%f32" for float32 printed as "%f"), formatted on the
    device (olv_write_text).  values: a DeviceBuffer or a numpy array."""
    from ._lib import OLV_PTR_HOST as H
    code = _TXT[fmt]
    if isinstance(values, DeviceBuffer):
        ptr, count, flags = values.ptr, values.nbytes // (stride or values.dtype.itemsize), 0
    else:
        a = np.ascontiguousarray(values)
        ptr, count, flags, values = a.ctypes.data, a.nbytes // (stride or a.dtype.itemsize), H, a
    if profile:
        flags |= OLV_PROFILE
    check(lib().olv_write_text(str(path).encode(), ptr, count, stride, code, flags,
                               ctypes.byref(stats) if stats is not None else None), "olv_write_text")


def write_binary(path, values, dtype_tag: int = 0):
    """Binary dump of a DeviceBuffer or numpy array (olv_write_binary)."""
    from ._lib import OLV_PTR_HOST as H
    if isinstance(values, DeviceBuffer):
        check(lib().olv_write_binary(str(path).encode(), values.ptr, values.count, values.dtype.itemsize, dtype_tag, 0),
              "olv_write_binary")
    else:
        a = np.ascontiguousarray(values)
        check(lib().olv_write_binary(str(path).encode(), a.ctypes.data, a.size, a.dtype.itemsize, dtype_tag, H),
              "olv_write_binary")


def read_binary(path, dtype) -> "DeviceBuffer":
    """A dump written by write_binary, read into device memory."""
    count, esz, tag = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_int()
    check(lib().olv_read_binary(str(path).encode(), None, 0, ctypes.byref(count), ctypes.byref(esz), ctypes.byref(tag),
                                0), "olv_read_binary")
    dt = np.dtype(dtype)
    if esz.value != dt.itemsize:
        raise ValueError(f"{path}: elements of {esz.value} bytes, asked for {dt}")
    out = DeviceBuffer(count.value, dt)
    check(lib().olv_read_binary(str(path).encode(), out.ptr, out.nbytes, None, None, None, 0), "olv_read_binary")
    return out


# ---------------------------------------------------------------- loaders
def load_edge_list(path) -> CsrGraph:
    """Edge-list file (csrGraph.h:106-171 format: '#' comments, ``n m``
    header, ``src dst [w]`` lines).  Unlike the reference, input need not be
    sorted (tuples are stable-sorted by source, the sort the reference leaves
    commented out at csrGraph.h:149) and a third column is kept as int32
    weights (dropped by the reference at csrGraph.h:139)."""
    header = None
    src, dst, wts = [], [], []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s[0] == "#":
                continue
            f = s.split()
            if header is None:
                header = (int(f[0]), int(f[1]))
                continue
            src.append(int(f[0]))
            dst.append(int(f[1]))
            wts.append(int(f[2]) if len(f) > 2 else None)
    if header is None:
        raise OliveError(1, f"{path}: no 'n m' header line")
    n, m = header
    if len(src) != m:
        raise OliveError(1, f"{path}: header says {m} edges, found {len(src)}")
    s = np.asarray(src, dtype=np.int64)
    d = np.asarray(dst, dtype=np.int64)
    if m and (s.min() < 0 or d.min() < 0 or s.max() >= n or d.max() >= n):
        raise OliveError(1, f"{path}: vertex id out of range [0, {n})")
    order = np.argsort(s, kind="stable")
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, s + 1, 1)
    w = None
    if m and all(x is not None for x in wts):
        w = np.asarray(wts, dtype=np.int32)[order]
    return CsrGraph(np.cumsum(rp).astype(np.uint32), d[order].astype(np.uint32), w)


def load_dimacs(path) -> CsrGraph:
    """METIS/DIMACS adjacency file (csrGraph.h:177-236): header ``n m``
    (m undirected edges), then line i lists the 1-based neighbours of
    vertex i-1."""
    rows = []
    header = None
    with open(path) as fh:
        for line in fh:
            if line.lstrip().startswith("#"):
                continue
            if header is None:
                if line.strip():
                    f = line.split()
                    header = (int(f[0]), int(f[1]))
                continue
            rows.append([int(x) - 1 for x in line.split()])
    n, m = header
    rows = (rows + [[]] * n)[:n]
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.asarray([x for r in rows for x in r], dtype=np.uint32)
    if col.size != 2 * m:
        raise OliveError(1, f"{path}: expected {2 * m} arcs, found {col.size}")
    return CsrGraph(rp.astype(np.uint32), col)


class Graph:
    """Device-resident graph handle (olv_graph): the reference's
    Oliver::readGraph (oliver.h:190-202) upload, plus weights and the in-edge
    CSR."""

    def __init__(self, csr: CsrGraph, transpose: bool = True):
        init()
        self.csr = csr
        self.n, self.m = csr.n, csr.m
        h = ctypes.c_void_p()
        wd = OLV_W_NONE
        if csr.weights is not None:
            wd = OLV_W_I32 if csr.weights.dtype == np.int32 else OLV_W_F32
        flags = OLV_PTR_HOST | (OLV_BUILD_TRANSPOSE if transpose else 0)
        check(lib().olv_graph_create(ctypes.byref(h), csr.n, csr.m, csr.row_ptr.ctypes.data,
                                     _np_ptr(csr.col), _np_ptr(csr.weights), wd, flags), "olv_graph_create")
        self.h = h
        _live.add(self)

    @classmethod
    def _adopt(cls, h, csr: CsrGraph | None = None):
        """Wrap a handle made by the library (device ingest); the host CSR is
        downloaded unless given."""
        g = cls.__new__(cls)
        g.h = h
        _live.add(g)
        n, m = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().olv_graph_counts(h, ctypes.byref(n), ctypes.byref(m)), "olv_graph_counts")
        g.n, g.m = n.value, m.value
        g.csr = csr if csr is not None else g.to_host_csr()
        return g

    def to_host_csr(self) -> CsrGraph:
        """The resident out-CSR (and weights) copied to the host."""
        rp = np.empty(self.n + 1, np.uint32)
        col = np.empty(max(self.m, 1), np.uint32)
        w = np.empty(max(self.m, 1), np.uint32)
        wd = ctypes.c_int()
        check(lib().olv_graph_csr_to_host(self.h, rp.ctypes.data, col.ctypes.data, w.ctypes.data, ctypes.byref(wd)),
              "olv_graph_csr_to_host")
        wt = None
        if wd.value == OLV_W_I32:
            wt = w[:self.m].view(np.int32).copy()
        elif wd.value == OLV_W_F32:
            wt = w[:self.m].view(np.float32).copy()
        return CsrGraph(rp, col[:self.m].copy(), wt)

    def transpose_to_host(self):
        """(in_row_ptr, in_col, in_eid) of the resident in-CSR."""
        irp = np.empty(self.n + 1, np.uint32)
        icol = np.empty(max(self.m, 1), np.uint32)
        ieid = np.empty(max(self.m, 1), np.uint32)
        check(lib().olv_graph_transpose_to_host(self.h, irp.ctypes.data, icol.ctypes.data, ieid.ctypes.data),
              "olv_graph_transpose_to_host")
        return irp, icol[:self.m].copy(), ieid[:self.m].copy()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            if _native_free_allowed():
                lib().olv_graph_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ BFS/SSSP
    def bfs(self, src: int, out: DeviceBuffer | None = None, flags: int = 0):
        out = out or DeviceBuffer(self.n, np.int32)
        st = Stats()
        check(lib().olv_bfs(self.h, src, out.c, flags, ctypes.byref(st)), "olv_bfs")
        return out, st

    def bfs_multisource(self, srcs, out: DeviceBuffer | None = None, flags: int = 0):
        s = np.ascontiguousarray(srcs, dtype=np.uint32)
        out = out or DeviceBuffer(self.n, np.int32)
        st = Stats()
        check(lib().olv_bfs_multisource(self.h, s.ctypes.data, s.size, out.c, flags | 0, ctypes.byref(st)),
              "olv_bfs_multisource")
        return out, st

    def sssp(self, src: int, weights: DeviceBuffer | None = None, out: DeviceBuffer | None = None, flags: int = 0):
        out = out or DeviceBuffer(self.n, np.int32)
        st = Stats()
        wp = weights.c if weights is not None else None
        check(lib().olv_sssp(self.h, src, wp, out.c, flags, ctypes.byref(st)), "olv_sssp")
        return out, st

    def bfs_batched(self, srcs, out: DeviceBuffer | None = None, flags: int = 0):
        s = np.ascontiguousarray(srcs, dtype=np.uint32)
        out = out or DeviceBuffer(self.n * s.size, np.int32)
        st = Stats()
        check(lib().olv_bfs_batched(self.h, s.ctypes.data, s.size, out.c, flags | 0, ctypes.byref(st)),
              "olv_bfs_batched")
        return out, st

    def sssp_batched(self, srcs, weights: DeviceBuffer | None = None, out: DeviceBuffer | None = None,
                     flags: int = 0):
        s = np.ascontiguousarray(srcs, dtype=np.uint32)
        out = out or DeviceBuffer(self.n * s.size, np.int32)
        st = Stats()
        wp = weights.c if weights is not None else None
        check(lib().olv_sssp_batched(self.h, s.ctypes.data, s.size, wp, out.c, flags, ctypes.byref(st)),
              "olv_sssp_batched")
        return out, st

    # ------------------------------------------------------------- PageRank
    def pagerank(self, iters: int = 20, damping: float = 0.85, epsilon: float = 1e-7,
                 out: DeviceBuffer | None = None, flags: int = 0):
        out = out or DeviceBuffer(self.n, np.float32)
        st = Stats()
        check(lib().olv_pagerank_f32(self.h, damping, iters, epsilon, out.c, flags, ctypes.byref(st)),
              "olv_pagerank_f32")
        return out, st

    def pagerank_weighted(self, w: DeviceBuffer, t: DeviceBuffer, damping=0.85, tol=1e-6, max_iters=1000,
                          out: DeviceBuffer | None = None, flags: int = 0):
        out = out or DeviceBuffer(self.n, np.float32)
        st = Stats()
        check(lib().olv_pagerank_weighted_f32(self.h, w.c, t.c, damping, tol, max_iters, out.c, flags,
                                              ctypes.byref(st)), "olv_pagerank_weighted_f32")
        return out, st

    def pagerank_vjp(self, w: DeviceBuffer, r: DeviceBuffer, gbar: DeviceBuffer, damping=0.85, tol=1e-6,
                     max_iters=1000, flags: int = 0):
        gw = DeviceBuffer(self.m, np.float32)
        gt = DeviceBuffer(self.n, np.float32)
        lam = DeviceBuffer(self.n, np.float32)
        st = Stats()
        check(lib().olv_pagerank_vjp_f32(self.h, w.c, r.c, gbar.c, damping, tol, max_iters, gw.c, gt.c, lam.c, flags,
                                         ctypes.byref(st)), "olv_pagerank_vjp_f32")
        return gw, gt, lam, st

    def pagerank_jvp(self, w: DeviceBuffer, r: DeviceBuffer, dw: DeviceBuffer, dt: DeviceBuffer, damping=0.85,
                     tol=1e-6, max_iters=1000, flags: int = 0):
        """Tangent dr of the weighted fixed point for (dw, dt) at r
        (olv_pagerank_jvp_f32)."""
        dr = DeviceBuffer(self.n, np.float32)
        st = Stats()
        check(lib().olv_pagerank_jvp_f32(self.h, w.c, r.c, dw.c, dt.c, damping, tol, max_iters, dr.c, flags,
                                         ctypes.byref(st)), "olv_pagerank_jvp_f32")
        return dr, st

    def transpose_host(self):
        rp = np.empty(self.n + 1, np.uint32)
        ci = np.empty(max(self.m, 1), np.uint32)
        ei = np.empty(max(self.m, 1), np.uint32)
        check(lib().olv_graph_transpose_to_host(self.h, rp.ctypes.data, ci.ctypes.data, ei.ctypes.data),
              "olv_graph_transpose_to_host")
        return rp, ci[:self.m], ei[:self.m]

    def autotune(self):
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().olv_autotune(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "olv_autotune")
        return a.value, b.value, c.value

    def tuning(self) -> Tuning:
        """The handle's current olv_tuning."""
        t = Tuning()
        check(lib().olv_get_tuning_ex(self.h, ctypes.byref(t)), "olv_get_tuning_ex")
        return t

    def set_tuning(self, t: Tuning):
        check(lib().olv_set_tuning_ex(self.h, ctypes.byref(t)), "olv_set_tuning_ex")

    def autotune_ex(self, what: int = 0xF, cache_path: str | None = None) -> Tuning:
        """olv_autotune_ex: search the kernel families in `what` on this graph
        (or take the cached result for its signature from cache_path)."""
        t = Tuning()
        check(lib().olv_autotune_ex(self.h, what, None if cache_path is None else str(cache_path).encode(),
                                    ctypes.byref(t)), "olv_autotune_ex")
        return t

    def signature(self) -> str:
        buf = ctypes.create_string_buffer(96)
        check(lib().olv_graph_signature(self.h, buf, 96), "olv_graph_signature")
        return buf.value.decode()

    def set_buckets(self, thread_max, wave_max, block_threads=0):
        check(lib().olv_set_buckets(self.h, thread_max, wave_max, block_threads), "olv_set_buckets")
