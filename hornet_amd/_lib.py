"""ctypes binding of libhornet_hip.so (declarations in include/hornet_hip.h).

The library is built in-tree by ``hornet_amd/csrc/Makefile`` (``build()`` in
``__graft_entry__``).  Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhornet_hip.so")

HN_OK = 0
HN_E_INVALID_ARG = 1
HN_E_OOM = 2
HN_E_HIP = 3
HN_E_NO_DEVICE = 4
HN_E_CAPACITY = 5

HN_DEVICE_PTRS = 0x1
HN_OFFSETS_INT32 = 0x2
HN_DEDUP_BATCH = 0x4
HN_DEDUP_GRAPH = 0x8
HN_DROP_SELF_LOOPS = 0x10
HN_SYMMETRIZE = 0x20

# StructureProp / ParsingProp bits of the graph-file reader
HN_GS_DIRECTED, HN_GS_UNDIRECTED, HN_GS_ENABLE_INGOING = 0x1, 0x2, 0x4
HN_PARSE_RANDOMIZE, HN_PARSE_SORT, HN_PARSE_PRINT_INFO = 0x1, 0x2, 0x4
HN_PARSE_RM_SINGLETON, HN_PARSE_DIRECTED_BY_DEGREE = 0x8, 0x10


class HornetError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


class HnVrec(ctypes.Structure):
    _fields_ = [("off", ctypes.c_int64), ("deg", ctypes.c_int32), ("cap_log2", ctypes.c_int32)]


HN_MAX_EDGE_META = 4


class HnEdgeMeta(ctypes.Structure):
    _fields_ = [("nfields", ctypes.c_int32), ("bytes", ctypes.c_int32 * HN_MAX_EDGE_META)]


class HnDview(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int32), ("nmeta", ctypes.c_int32), ("ne", ctypes.c_int64),
                ("rec", ctypes.c_void_p), ("arena", ctypes.c_void_p),
                ("meta", ctypes.c_void_p * HN_MAX_EDGE_META),
                ("meta_bytes", ctypes.c_int32 * HN_MAX_EDGE_META)]


class HnBfsStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("time_kernels", ctypes.c_int32),
                ("visited", ctypes.c_int64), ("edges_scanned", ctypes.c_int64),
                ("expand_ns", ctypes.c_int64), ("expand_launches", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class HnUpdateStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("applied", ctypes.c_int64), ("touched", ctypes.c_int64),
                ("relocated", ctypes.c_int64), ("touched_deg", ctypes.c_int64), ("relocated_deg", ctypes.c_int64),
                ("moved", ctypes.c_int64), ("kind", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class HnMsbfsStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("push_levels", ctypes.c_int32),
                ("pull_levels", ctypes.c_int32), ("pad_", ctypes.c_int32), ("edges", ctypes.c_int64)]


class HnRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("source", ctypes.c_int32), ("n", ctypes.c_int64),
                ("src", ctypes.c_void_p), ("dst", ctypes.c_void_p)]


class HnResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("kind", ctypes.c_int32), ("status", ctypes.c_int32),
                ("t_arrive_ns", ctypes.c_int64), ("t_start_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64),
                ("reached", ctypes.c_int64), ("dist_sum", ctypes.c_int64), ("max_depth", ctypes.c_int32),
                ("batch", ctypes.c_int32), ("applied", ctypes.c_int64), ("version", ctypes.c_uint64)]


class HnServeConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int32), ("slots", ctypes.c_int32), ("mem_budget_bytes", ctypes.c_int64),
                ("plain", ctypes.c_int32), ("min_window_queries", ctypes.c_int32)]


HN_REQ_BFS, HN_REQ_TWO_HOP, HN_REQ_INSERT, HN_REQ_ERASE = 0, 1, 2, 3

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
D = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)
PU64 = ctypes.POINTER(ctypes.c_uint64)

# name -> argtypes (restype is hn_status unless listed in _RESTYPES)
SIGNATURES = {
    "hn_status_string": [I32],
    "hn_last_error": [],
    "hn_version": [],
    "hn_device_info": [PI32, PI64, PI64, PI32],
    "hn_graph_create_csr": [I32, I64, P, P, U32, D, P, PP],
    "hn_graph_create_coo": [I32, I64, P, P, U32, D, P, PP],
    "hn_graph_create_csr_meta": [I32, I64, P, P, ctypes.POINTER(HnEdgeMeta), P, U32, D, P, PP],
    "hn_graph_edge_meta": [P, ctypes.POINTER(HnEdgeMeta)],
    "hn_graph_export_coo_meta": [P, P, P, P, I64, PI64, P],
    "hn_graph_insert_meta": [P, P, P, P, I64, U32, PI64, P],
    "hn_graph_destroy": [P],
    "hn_graph_nv": [P, PI32],
    "hn_graph_ne": [P, PI64],
    "hn_graph_device_view": [P, ctypes.POINTER(HnDview)],
    "hn_graph_memory": [P, PI64, PI64, P],
    "hn_graph_export_coo": [P, P, P, I64, PI64, P],
    "hn_graph_export_csr": [P, P, P, I64, P],
    "hn_graph_export_csr_meta": [P, P, P, P, I64, P],
    "hn_graph_degrees": [P, P, P],
    "hn_graph_snapshot": [P, D, P, PP],
    "hn_graph_max_degree": [P, PI32, PI32, P],
    "hn_graph_insert": [P, P, P, I64, U32, PI64, P],
    "hn_graph_erase": [P, P, P, I64, U32, PI64, P],
    "hn_graph_update_stats": [P, ctypes.POINTER(HnUpdateStats), P],
    "hn_bfs_topdown": [P, I32, P, ctypes.POINTER(HnBfsStats), P],
    "hn_bfs_topdown_noalloc": [P, I32, P, ctypes.POINTER(HnBfsStats), P],
    "hn_bfs_topdown_noalloc_limit": [PI64],
    "hn_bfs_multi": [P, P, I32, P, ctypes.POINTER(HnMsbfsStats), P],
    "hn_bfs_multi_depth": [P, P, P, I32, P, ctypes.POINTER(HnMsbfsStats), P],
    "hn_bfs_multi_params": [P, I32, I32],
    "hn_bfs_multi_tune": [P, P, I32, P, PI32, PI32, P],
    "hn_bfs_multi_lb_tile": [P, I32],
    "hn_bfs_multi_get_params": [P, PI32, PI32, PI32],
    "hn_tc_count": [P, PU64, P, P],
    "hn_tc2_stored": [P, PU64, P, P],
    "hn_gen_rmat_pairs": [I32, I64, I64, U64, P, P, P],
    "hn_gen_rmat_graph": [I32, I32, U64, D, P, PP],
    "hn_gen_uniform_pairs": [I32, I64, U64, P, P, P],
    "hn_server_create": [P, ctypes.POINTER(HnServeConfig), PP],
    "hn_server_submit": [P, ctypes.POINTER(HnRequest), PI64],
    "hn_server_replay": [P, ctypes.POINTER(HnRequest), PI64, I64],
    "hn_server_drain": [P],
    "hn_server_results": [P, ctypes.POINTER(HnResult), I64, PI64],
    "hn_server_slots": [P, PI32],
    "hn_server_slot_bytes": [P, I32, PI64],
    "hn_server_destroy": [P],
    "hn_gen_chung_lu_pairs": [I32, D, I64, I64, U64, P, P, P],
    "hn_gen_chung_lu_graph": [I32, I32, D, U64, D, P, PP],
    "hn_gstd_read": [ctypes.c_char_p, U32, U32, U64, PP],
    "hn_gstd_destroy": [P],
    "hn_gstd_from_csr": [I32, I64, P, P, PP],
    "hn_gstd_info": [P, PI32, PI64, ctypes.POINTER(ctypes.c_uint32)],
    "hn_gstd_csr_out": [P, PP, PP],
    "hn_gstd_csr_in": [P, PP, PP],
    "hn_gstd_max_out_degree_id": [P, PI32],
    "hn_gstd_write_binary": [P, ctypes.c_char_p, I32],
    "hn_gstd_write_market": [P, ctypes.c_char_p],
}
_RESTYPES = {"hn_status_string": ctypes.c_char_p, "hn_last_error": ctypes.c_char_p,
             "hn_version": ctypes.c_char_p}

_lib = None


def header_symbols(path: str | None = None) -> list[str]:
    """Function names declared in include/hornet_hip.h."""
    import re
    if path is None:
        path = os.path.join(os.path.dirname(_HERE), "include", "hornet_hip.h")
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hn_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C hornet_amd/csrc` "
            "(or __graft_entry__.build()); there is no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ctypes.c_int32)
    _lib = lib
    return lib


def check(status: int, where: str):
    if status != HN_OK:
        lib = load()
        msg = lib.hn_last_error().decode(errors="replace")
        raise HornetError(status, where, msg)


def call(name: str, *args):
    lib = load()
    check(getattr(lib, name)(*args), name)
