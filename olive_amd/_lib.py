"""ctypes binding of libolive_hip.so (declarations in include/olive_hip.h).

The library must have been built (``python -m olive_amd.build`` or
``__graft_entry__.build()``); there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# OLIVE_LIB selects another build of the same sources (e.g. one compiled with
# different tuning macros by `python -m olive_amd.build --define ... --out`).
LIB_PATH = os.environ.get("OLIVE_LIB") or os.path.join(HERE, "libolive_hip.so")

OLV_OK = 0
OLV_ENOCONV = 6
STATUS = {0: "OLV_OK", 1: "OLV_EINVAL", 2: "OLV_ENOMEM", 3: "OLV_EHIP", 4: "OLV_ERANGE",
          5: "OLV_EUNSUPPORTED", 6: "OLV_ENOCONV", 7: "OLV_EINTERNAL"}

OLV_PTR_HOST = 0x1
OLV_NO_DIRECTION_OPT = 0x2
OLV_FORCE_BOTTOM_UP = 0x4
OLV_HOST_LOOP = 0x8
OLV_PROFILE = 0x10
OLV_PERSISTENT = 0x20
OLV_GRAPH_ORDER = 0x40
OLV_BUILD_TRANSPOSE = 0x100
OLV_W_UNCHANGED = 0x200
OLV_PR_ROW_PULL = 0x400
OLV_SSSP_DIRECT = 0x800
OLV_TUNE_BFS, OLV_TUNE_BATCHED, OLV_TUNE_PULL, OLV_TUNE_SSSP, OLV_TUNE_ALL = 0x1, 0x2, 0x4, 0x8, 0xF

OLV_W_NONE, OLV_W_I32, OLV_W_F32 = 0, 1, 2
OLV_FMT_EDGE_LIST, OLV_FMT_DIMACS = 0, 1
OLV_TXT_I32, OLV_TXT_F32, OLV_TXT_F64 = 0, 1, 2


class OliveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("elapsed_ms", ctypes.c_double),
        ("edges_traversed", ctypes.c_uint64),
        ("vertices_reached", ctypes.c_uint64),
        ("edges_inspected", ctypes.c_uint64),
        ("iterations", ctypes.c_uint32),
        ("bottom_up_levels", ctypes.c_uint32),
        ("residual", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    """olv_tuning: launch shapes and thresholds of a graph handle."""
    _fields_ = [
        ("bfs_expand_per_cu", ctypes.c_uint32),
        ("bfs_expand_per_cu_twin", ctypes.c_uint32),
        ("bfs_compact_per_cu", ctypes.c_uint32),
        ("td_wave_thr", ctypes.c_uint32),
        ("td_block_thr", ctypes.c_uint32),
        ("hub_thr", ctypes.c_uint32),
        ("alpha", ctypes.c_float),
        ("beta", ctypes.c_float),
        ("mbfs_pull_per_cu", ctypes.c_uint32),
        ("seg_grid_per_cu", ctypes.c_uint32),
        ("seg_unroll", ctypes.c_uint32),
        ("seg_bounds", ctypes.c_uint32 * 3),
        ("sssp_relax_per_cu", ctypes.c_uint32),
        ("sssp_delta_num", ctypes.c_uint32),
        ("source", ctypes.c_uint32),
        ("tune_ms", ctypes.c_double),
    ]
    SOURCES = {0: "defaults", 1: "measured", 2: "cache", 3: "caller"}

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["seg_bounds"] = list(self.seg_bounds)
        d["source"] = self.SOURCES.get(self.source, self.source)
        return d


class HostCsr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("m", ctypes.c_uint32),
        ("row_ptr", ctypes.POINTER(ctypes.c_uint32)),
        ("col", ctypes.POINTER(ctypes.c_uint32)),
        ("weights", ctypes.c_void_p),
        ("weight_dtype", ctypes.c_int),
    ]


_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_sz = ctypes.c_size_t
_int = ctypes.c_int
_uint = ctypes.c_uint
_statp = ctypes.POINTER(Stats)
_csrp = ctypes.POINTER(HostCsr)

# name -> (restype, argtypes)
SIGNATURES = {
    "olv_last_error": (ctypes.c_char_p, []),
    "olv_version": (_int, []),
    "olv_init": (_int, [_int]),
    "olv_set_stream": (_int, [_vp]),
    "olv_get_stream": (_vp, []),
    "olv_synchronize": (_int, []),
    "olv_free_cached_memory": (_int, []),
    "olv_device_info": (_int, [ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_sz)]),
    "olv_malloc": (_int, [ctypes.POINTER(_vp), _sz]),
    "olv_free": (_int, [_vp]),
    "olv_host_malloc": (_int, [ctypes.POINTER(_vp), _sz]),
    "olv_host_free": (_int, [_vp]),
    "olv_memcpy_h2d": (_int, [_vp, _vp, _sz]),
    "olv_memcpy_d2h": (_int, [_vp, _vp, _sz]),
    "olv_memcpy_d2d": (_int, [_vp, _vp, _sz]),
    "olv_fill": (_int, [_vp, _u64, _sz, _int]),
    "olv_bitmap_words": (_int, [_u32, _u32p]),
    "olv_bitmap_to_queue": (_int, [_vp, _u32, _vp, _vp]),
    "olv_queue_to_bitmap": (_int, [_vp, _vp, _vp, _u32, _int]),
    "olv_frontier_bin": (_int, [_vp, _vp, _vp, ctypes.POINTER(_vp), _vp]),
    "olv_graph_buckets": (_int, [_vp, _u32p, _u32p]),
    "olv_frontier_split": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "olv_hub_capacity": (_sz, [_u64]),
    "olv_graph_create": (_int, [ctypes.POINTER(_vp), _u32, _u32, _vp, _vp, _vp, _int, _uint]),
    "olv_graph_destroy": (_int, [_vp]),
    "olv_graph_build_transpose": (_int, [_vp]),
    "olv_graph_counts": (_int, [_vp, _u32p, _u32p]),
    "olv_graph_device_arrays": (_int, [_vp] + [ctypes.POINTER(_vp)] * 6),
    "olv_graph_transpose_to_host": (_int, [_vp, _vp, _vp, _vp]),
    "olv_graph_from_coo": (_int, [ctypes.POINTER(_vp), _u32, _u32, _vp, _vp, _vp, _int, _uint, _statp]),
    "olv_graph_from_edge_list_text": (_int, [ctypes.POINTER(_vp), _vp, _sz, _uint, _statp]),
    "olv_graph_from_dimacs_text": (_int, [ctypes.POINTER(_vp), _vp, _sz, _uint, _statp]),
    "olv_graph_load_file": (_int, [ctypes.POINTER(_vp), ctypes.c_char_p, _int, _uint, _statp]),
    "olv_graph_csr_to_host": (_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(_int)]),
    "olv_write_text": (_int, [ctypes.c_char_p, _vp, _sz, _sz, _int, _uint, _statp]),
    "olv_write_binary": (_int, [ctypes.c_char_p, _vp, _sz, _sz, _int, _uint]),
    "olv_read_binary": (_int, [ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_u64), _u32p, ctypes.POINTER(_int), _uint]),
    "olv_bfs": (_int, [_vp, _u32, _vp, _uint, _statp]),
    "olv_bfs_noalloc": (_int, [_vp, _u32, _vp, _uint, _statp]),
    "olv_bfs_multisource": (_int, [_vp, _vp, _u32, _vp, _uint, _statp]),
    "olv_sssp": (_int, [_vp, _u32, _vp, _vp, _uint, _statp]),
    "olv_pagerank_f32": (_int, [_vp, _f32, _u32, _f64, _vp, _uint, _statp]),
    "olv_pagerank_weighted_f32": (_int, [_vp, _vp, _vp, _f32, _f32, _u32, _vp, _uint, _statp]),
    "olv_pagerank_vjp_f32": (_int, [_vp, _vp, _vp, _vp, _f32, _f32, _u32, _vp, _vp, _vp, _uint, _statp]),
    "olv_pagerank_jvp_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _f32, _f32, _u32, _vp, _uint, _statp]),
    "olv_bfs_batched": (_int, [_vp, _vp, _u32, _vp, _uint, _statp]),
    "olv_sssp_batched": (_int, [_vp, _vp, _u32, _vp, _vp, _uint, _statp]),
    "olv_bfs_multigraph": (_int, [ctypes.POINTER(_vp), _u32, _vp, ctypes.POINTER(_vp), _uint, _statp]),
    "olv_graph_union": (_int, [ctypes.POINTER(_vp), _u32, ctypes.POINTER(_vp), _vp]),
    "olv_plan_bfs": (_int, [_vp, _vp, _u32, _vp, _uint, ctypes.POINTER(_vp)]),
    "olv_plan_bfs_batched": (_int, [_vp, _vp, _u32, _vp, _uint, ctypes.POINTER(_vp)]),
    "olv_plan_launch": (_int, [_vp, _statp]),
    "olv_plan_destroy": (_int, [_vp]),
    "olv_autotune": (_int, [_vp, _u32p, _u32p, _u32p]),
    "olv_set_buckets": (_int, [_vp, _u32, _u32, _u32]),
    "olv_set_direction": (_int, [_vp, _f32, _f32]),
    "olv_get_tuning": (_int, [_vp, _u32p, _u32p, ctypes.POINTER(_f32), ctypes.POINTER(_f32)]),
    "olv_get_tuning_ex": (_int, [_vp, ctypes.POINTER(Tuning)]),
    "olv_set_tuning_ex": (_int, [_vp, ctypes.POINTER(Tuning)]),
    "olv_autotune_ex": (_int, [_vp, _uint, ctypes.c_char_p, ctypes.POINTER(Tuning)]),
    "olv_graph_signature": (_int, [_vp, ctypes.c_char_p, _sz]),
    "olv_csr_signature": (_int, [_u32, _u32, _vp, _vp, ctypes.c_char_p, _sz]),
    "olv_tuning_cache_load": (_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Tuning), ctypes.POINTER(_int)]),
    "olv_tuning_cache_store": (_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Tuning)]),
    "olv_gen_grid": (_int, [_u32, _u32, _csrp]),
    "olv_gen_rmat": (_int, [_u32, _u32, _f64, _f64, _f64, _u64, _int, _csrp]),
    "olv_gen_uniform": (_int, [_u32, _u32, _u64, _csrp]),
    "olv_gen_weights_i32": (_int, [_csrp, _i32, _i32, _u64]),
    "olv_gen_weights_f32_normalized": (_int, [_csrp, _u64]),
    "olv_gen_dirichlet": (_int, [_u32, _u64, _vp]),
    "olv_gen_batch_scale": (_u32, [_u64, _u32, _u32, _u32]),
    "olv_gen_batch_seed": (_u64, [_u64, _u32]),
    "olv_host_csr_free": (None, [_csrp]),
}

_lib = None


def lib():
    """Load libolive_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OliveError(5, f"{LIB_PATH} is missing: build it with `python -m olive_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code, what=""):
    if code != OLV_OK:
        msg = lib().olv_last_error().decode(errors="replace")
        raise OliveError(code, f"{what}: {msg}" if what else msg)
    return code
