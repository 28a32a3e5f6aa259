"""Host-side mirror of the reference's Hornet API over the C ABI.

Names, argument meaning and error behaviour follow the reference:

* ``HornetInit(nV, nE, offsets, edges)``  -- hornet/include/Core/HornetInitialize/HornetInit.cuh:70-73
* ``Hornet(init)``                        -- hornet/include/Core/Hornet.cuh:105-129
  (``nV``, ``nE``, ``insert``, ``erase``, ``getCOO``, ``max_degree_id``,
  ``max_degree``, ``device``)
* ``BatchUpdatePtr(n, src, dst)`` / ``BatchUpdate`` -- hornet/include/Core/BatchUpdate/BatchUpdate.cuh:95-197
* ``BfsTopDown``                          -- hornetsnest/include/Static/BreadthFirstSearch/TopDown.cuh:52-70
* ``TriangleCounting2``                   -- hornetsnest/include/Static/TriangleCounting/triangle2.cuh:18-37
* ``GraphStd(structure).read(path, parsing)`` -- xlib/include/Graph/GraphStd.hpp,
  xlib/src/Graph/GraphBase.cpp:130-206 (host graph files; no device needed)

Device memory is held in torch tensors (plumbing only); every computation runs
in libhornet_hip.so.  Where the reference exits the process on an error, these
raise ``HornetError``.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HornetError, call

INF = np.iinfo(np.int32).max


def _dev() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("hornet_amd needs a HIP device (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _as_device_i32(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32))
    return t.to(device=_dev(), dtype=torch.int32).contiguous()


def _meta_arrays(meta, n, where):
    """Edge meta-data fields as contiguous 4- or 8-byte numpy arrays of n elements."""
    out = []
    for m in meta or ():
        a = np.ascontiguousarray(m)
        if a.dtype.itemsize not in (4, 8) or a.ndim != 1 or a.size != n:
            raise HornetError(_lib.HN_E_INVALID_ARG, where,
                              "meta-data fields are 1-d arrays of 4- or 8-byte elements, one per edge")
        out.append(a)
    if len(out) > _lib.HN_MAX_EDGE_META:
        raise HornetError(_lib.HN_E_INVALID_ARG, where, "more than HN_MAX_EDGE_META meta-data fields")
    return out


def _ptr_array(ptrs):
    arr = (ctypes.c_void_p * _lib.HN_MAX_EDGE_META)()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


class HornetInit:
    """Host CSR (offsets[nV+1], edges[nE]) used to build a Hornet, with
    optional edge meta-data (HornetInit::insertEdgeData, HornetInit.cuh:75-97)."""

    def __init__(self, nV: int, nE: int, offsets, edges):
        self.nV = int(nV)
        self.nE = int(nE)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.edges = np.ascontiguousarray(edges, dtype=np.int32)
        self.meta = []
        if self.offsets.size != self.nV + 1 or self.edges.size != self.nE:
            raise HornetError(_lib.HN_E_INVALID_ARG, "HornetInit", "size mismatch")

    def insertEdgeData(self, *fields):
        """One array per meta-data field, nE elements each (4 or 8 bytes)."""
        self.meta = _meta_arrays(fields, self.nE, "HornetInit.insertEdgeData")
        return self


class structure_prop:
    """StructureProp bits (xlib/include/Graph/GraphBase.hpp:92-93,115-123)."""
    NONE = 0
    DIRECTED = _lib.HN_GS_DIRECTED
    UNDIRECTED = _lib.HN_GS_UNDIRECTED
    ENABLE_INGOING = _lib.HN_GS_ENABLE_INGOING


class parsing_prop:
    """ParsingProp bits (GraphBase.hpp:46-48); PRINT_INFO prints nothing."""
    NONE = 0
    RANDOMIZE = _lib.HN_PARSE_RANDOMIZE
    SORT = _lib.HN_PARSE_SORT
    PRINT_INFO = _lib.HN_PARSE_PRINT_INFO
    RM_SINGLETON = _lib.HN_PARSE_RM_SINGLETON
    DIRECTED_BY_DEGREE = _lib.HN_PARSE_DIRECTED_BY_DEGREE


class GraphStd:
    """Host graph read from a MatrixMarket (.mtx), DIMACS9 (.gr) or binary
    (.bin) file: ``graph::GraphStd`` (xlib/include/Graph/GraphStd.hpp).

    ``GraphStd(structure)`` then ``read(filename, parsing)`` as in the
    reference drivers (hornetsnest/test/TriangleTest2.cu:86-87); the CSR
    accessors return numpy views valid while the object lives.  Parsing and
    the COO -> CSR pass run in libhornet_hip.so on the host.  RANDOMIZE takes
    an explicit ``seed`` (the reference seeds from the clock).
    """

    def __init__(self, structure: int = structure_prop.NONE):
        self._structure = int(structure)
        self._h = None

    def read(self, filename, parsing: int = parsing_prop.NONE, seed: int = 0) -> "GraphStd":
        self._free()
        h = ctypes.c_void_p()
        call("hn_gstd_read", str(filename).encode(), self._structure, int(parsing), int(seed),
             ctypes.byref(h))
        self._h = h
        return self

    @classmethod
    def from_csr(cls, offsets, edges) -> "GraphStd":
        """GraphStd(csr_offsets, nV, csr_edges, nE) (xlib/src/Graph/GraphStd.cpp:50-59):
        an UNDIRECTED host graph copied from a CSR (e.g. Hornet.getCSR()),
        so it can be written with writeBinary and read back later."""
        off = np.ascontiguousarray(offsets, np.int64)
        adj = np.ascontiguousarray(edges, np.int32)
        g = cls(structure_prop.UNDIRECTED)
        h = ctypes.c_void_p()
        call("hn_gstd_from_csr", off.size - 1, adj.size, off.ctypes.data_as(ctypes.c_void_p),
             adj.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
        g._h = h
        return g

    def _info(self):
        if self._h is None:
            raise HornetError(_lib.HN_E_INVALID_ARG, "GraphStd", "no graph read")
        nv, ne, st = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_uint32()
        call("hn_gstd_info", self._h, ctypes.byref(nv), ctypes.byref(ne), ctypes.byref(st))
        return nv.value, ne.value, st.value

    def nV(self) -> int:
        return self._info()[0]

    def nE(self) -> int:
        return self._info()[1]

    def is_directed(self) -> bool:
        return bool(self._info()[2] & structure_prop.DIRECTED)

    def is_undirected(self) -> bool:
        return bool(self._info()[2] & structure_prop.UNDIRECTED)

    def _csr(self, name):
        nv, ne, _ = self._info()
        o, e = ctypes.c_void_p(), ctypes.c_void_p()
        call(name, self._h, ctypes.byref(o), ctypes.byref(e))
        off = np.ctypeslib.as_array(ctypes.cast(o, ctypes.POINTER(ctypes.c_int64)), (nv + 1,))
        adj = (np.ctypeslib.as_array(ctypes.cast(e, ctypes.POINTER(ctypes.c_int32)), (ne,))
               if ne > 0 else np.empty(0, np.int32))
        return off, adj

    def csr_out_offsets(self) -> np.ndarray:
        return self._csr("hn_gstd_csr_out")[0]

    def csr_out_edges(self) -> np.ndarray:
        return self._csr("hn_gstd_csr_out")[1]

    def csr_in_offsets(self) -> np.ndarray:
        return self._csr("hn_gstd_csr_in")[0]

    def csr_in_edges(self) -> np.ndarray:
        return self._csr("hn_gstd_csr_in")[1]

    def out_degrees(self) -> np.ndarray:
        return np.diff(self.csr_out_offsets()).astype(np.int32)

    def max_out_degree_id(self) -> int:
        self._info()
        v = ctypes.c_int32()
        call("hn_gstd_max_out_degree_id", self._h, ctypes.byref(v))
        return v.value

    def writeBinary(self, filename, offset_bytes: int = 4):
        self._info()
        call("hn_gstd_write_binary", self._h, str(filename).encode(), int(offset_bytes))

    def writeMarket(self, filename):
        self._info()
        call("hn_gstd_write_market", self._h, str(filename).encode())

    def hornet_init(self) -> "HornetInit":
        """HornetInit(graph.nV(), graph.nE(), graph.csr_out_offsets(),
        graph.csr_out_edges()) -- hornetsnest/test/BFSTest.cu:22-23."""
        off, adj = self._csr("hn_gstd_csr_out")
        init = HornetInit(off.size - 1, adj.size, off, adj)
        init._owner = self   # the arrays view this graph's memory
        return init

    def _free(self):
        if self._h is not None:
            _lib.load().hn_gstd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class BatchUpdatePtr:
    """(n, src, dst) batch; host arrays or device tensors."""

    def __init__(self, n: int, src, dst):
        self.n = int(n)
        self.src = src
        self.dst = dst


class BatchUpdate:
    """A batch copied to the device (BatchUpdate::reset copies it,
    hornet/include/Core/BatchUpdate/BatchUpdate.i.cuh:154-184)."""

    def __init__(self, ptr_or_src, dst=None, meta=None):
        if isinstance(ptr_or_src, BatchUpdatePtr):
            src, dst, n = ptr_or_src.src, ptr_or_src.dst, ptr_or_src.n
        else:
            src, n = ptr_or_src, len(ptr_or_src)
        self.src = _as_device_i32(src)[:n]
        self.dst = _as_device_i32(dst)[:n]
        if self.src.numel() != self.dst.numel():
            raise HornetError(_lib.HN_E_INVALID_ARG, "BatchUpdate", "src/dst length mismatch")
        # edge meta-data fields (one per graph field), copied to the device
        self.meta = [torch.from_numpy(a.copy()).to(_dev())
                     for a in _meta_arrays(meta, self.nE(), "BatchUpdate")] if meta else []

    def nE(self) -> int:
        return int(self.src.numel())

    size = nE


class Hornet:
    """hornet::gpu::Hornet<int> over an hn_graph handle."""

    def __init__(self, init: HornetInit | None = None, *, slack: float = 0.0, _handle=None):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        elif init is not None:
            layout = _lib.HnEdgeMeta()
            layout.nfields = len(init.meta)
            for i, a in enumerate(init.meta):
                layout.bytes[i] = a.dtype.itemsize
            ptrs = _ptr_array([a.ctypes.data for a in init.meta])
            call("hn_graph_create_csr_meta", init.nV, init.nE, init.offsets.ctypes.data_as(ctypes.c_void_p),
                 init.edges.ctypes.data_as(ctypes.c_void_p), ctypes.byref(layout), ptrs, 0, float(slack),
                 _stream(), ctypes.byref(self._h))
        else:
            raise HornetError(_lib.HN_E_INVALID_ARG, "Hornet", "need a HornetInit")

    # ---------------------------------------------------------------- factories
    @classmethod
    def from_csr(cls, offsets, edges, slack: float = 0.0) -> "Hornet":
        offsets = np.asarray(offsets)
        return cls(HornetInit(offsets.size - 1, int(offsets[-1]), offsets, edges), slack=slack)

    @classmethod
    def from_device_csr(cls, offsets: torch.Tensor, edges: torch.Tensor, meta=None,
                        slack: float = 0.0) -> "Hornet":
        """Hornet(HornetInit) from device tensors (int64 offsets, int32 edges)
        and optional device meta-data fields (4- or 8-byte dtypes, one
        element per edge)."""
        meta = list(meta or [])
        layout = _lib.HnEdgeMeta()
        layout.nfields = len(meta)
        for i, t in enumerate(meta):
            if t.element_size() not in (4, 8) or t.numel() != edges.numel() or not t.is_cuda:
                raise HornetError(_lib.HN_E_INVALID_ARG, "Hornet.from_device_csr",
                                  "meta-data fields are device tensors of 4- or 8-byte elements, one per edge")
            layout.bytes[i] = t.element_size()
        meta = [t.contiguous() for t in meta]
        off = offsets.to(torch.int64).contiguous()
        adj = edges.to(torch.int32).contiguous()
        h = ctypes.c_void_p()
        call("hn_graph_create_csr_meta", off.numel() - 1, adj.numel(), _ptr(off), _ptr(adj),
             ctypes.byref(layout), _ptr_array([t.data_ptr() for t in meta]), _lib.HN_DEVICE_PTRS,
             float(slack), _stream(), ctypes.byref(h))
        return cls(_handle=h)

    @classmethod
    def from_coo_device(cls, nv: int, src: torch.Tensor, dst: torch.Tensor, symmetrize=True,
                        drop_self_loops=True, slack: float = 0.0) -> "Hornet":
        h = ctypes.c_void_p()
        flags = _lib.HN_DEVICE_PTRS | (_lib.HN_SYMMETRIZE if symmetrize else 0) | \
            (_lib.HN_DROP_SELF_LOOPS if drop_self_loops else 0)
        call("hn_graph_create_coo", int(nv), int(src.numel()), _ptr(src), _ptr(dst), flags,
             float(slack), _stream(), ctypes.byref(h))
        return cls(_handle=h)

    @classmethod
    def rmat(cls, scale: int, seed: int, edge_factor: int = 16, slack: float = 0.0) -> "Hornet":
        h = ctypes.c_void_p()
        call("hn_gen_rmat_graph", int(scale), int(edge_factor), int(seed), float(slack), _stream(),
             ctypes.byref(h))
        return cls(_handle=h)

    @classmethod
    def chung_lu(cls, n: int, avg_degree: int, gamma: float, seed: int, slack: float = 0.0) -> "Hornet":
        h = ctypes.c_void_p()
        call("hn_gen_chung_lu_graph", int(n), int(avg_degree), float(gamma), int(seed), float(slack), _stream(),
             ctypes.byref(h))
        return cls(_handle=h)

    # ------------------------------------------------------------------ queries
    @property
    def handle(self):
        return self._h

    def nV(self) -> int:
        v = ctypes.c_int32()
        call("hn_graph_nv", self._h, ctypes.byref(v))
        return v.value

    def nE(self) -> int:
        v = ctypes.c_int64()
        call("hn_graph_ne", self._h, ctypes.byref(v))
        return v.value

    def device(self) -> _lib.HnDview:
        v = _lib.HnDview()
        call("hn_graph_device_view", self._h, ctypes.byref(v))
        return v

    def memory(self):
        used, cap = ctypes.c_int64(), ctypes.c_int64()
        call("hn_graph_memory", self._h, ctypes.byref(used), ctypes.byref(cap), _stream())
        return used.value, cap.value

    def max_degree_id(self) -> int:
        i, d = ctypes.c_int32(), ctypes.c_int32()
        call("hn_graph_max_degree", self._h, ctypes.byref(i), ctypes.byref(d), _stream())
        return i.value

    def max_degree(self) -> int:
        i, d = ctypes.c_int32(), ctypes.c_int32()
        call("hn_graph_max_degree", self._h, ctypes.byref(i), ctypes.byref(d), _stream())
        return d.value

    def degrees(self) -> torch.Tensor:
        out = torch.empty(max(self.nV(), 1), dtype=torch.int32, device=_dev())
        call("hn_graph_degrees", self._h, _ptr(out), _stream())
        return out[: self.nV()]

    def edge_meta_bytes(self) -> list:
        """Element size of each edge meta-data field ([] without meta-data)."""
        layout = _lib.HnEdgeMeta()
        call("hn_graph_edge_meta", self._h, ctypes.byref(layout))
        return [layout.bytes[i] for i in range(layout.nfields)]

    def getCOO(self, sort: bool = True, meta: bool = False):
        """(src, dst) device tensors sorted by (src, dst) (adjacency is kept sorted);
        with meta=True also a list of the meta-data fields (int32 / int64 tensors
        holding the fields' bits)."""
        ne = self.nE()
        src = torch.empty(max(ne, 1), dtype=torch.int32, device=_dev())
        dst = torch.empty(max(ne, 1), dtype=torch.int32, device=_dev())
        n = ctypes.c_int64()
        if not meta:
            call("hn_graph_export_coo", self._h, _ptr(src), _ptr(dst), ne, ctypes.byref(n), _stream())
            return src[:ne], dst[:ne]
        fields = [torch.empty(max(ne, 1), dtype=torch.int64 if b == 8 else torch.int32, device=_dev())
                  for b in self.edge_meta_bytes()]
        ptrs = _ptr_array([t.data_ptr() for t in fields])
        call("hn_graph_export_coo_meta", self._h, _ptr(src), _ptr(dst), ptrs, ne, ctypes.byref(n), _stream())
        return src[:ne], dst[:ne], [t[:ne] for t in fields]

    def getCSR(self):
        nv, ne = self.nV(), self.nE()
        off = torch.empty(nv + 1, dtype=torch.int64, device=_dev())
        adj = torch.empty(max(ne, 1), dtype=torch.int32, device=_dev())
        call("hn_graph_export_csr", self._h, _ptr(off), _ptr(adj), ne, _stream())
        return off, adj[:ne]

    def copy(self, slack: float = 2.0) -> "Hornet":
        """An independent, updatable device copy of the graph (hn_graph_snapshot)."""
        h = ctypes.c_void_p()
        call("hn_graph_snapshot", self._h, float(slack), _stream(), ctypes.byref(h))
        return Hornet(_handle=h)

    def snapshot(self, slack: float = 1.0) -> "HornetStatic":
        """A frozen device copy of the graph as it is now (hn_graph_snapshot):
        later updates of this graph do not reach it."""
        h = ctypes.c_void_p()
        call("hn_graph_snapshot", self._h, float(slack), _stream(), ctypes.byref(h))
        return HornetStatic(_handle=h)

    # ------------------------------------------------------------------ updates
    def insert(self, batch: BatchUpdate, removeBatchDuplicates: bool = False,
               removeGraphDuplicates: bool = False) -> int:
        flags = _lib.HN_DEVICE_PTRS
        if removeBatchDuplicates:
            flags |= _lib.HN_DEDUP_BATCH
        if removeGraphDuplicates:
            flags |= _lib.HN_DEDUP_GRAPH
        n = ctypes.c_int64()
        if batch.meta:
            ptrs = _ptr_array([t.data_ptr() for t in batch.meta])
            call("hn_graph_insert_meta", self._h, _ptr(batch.src), _ptr(batch.dst), ptrs, batch.nE(), flags,
                 ctypes.byref(n), _stream())
        else:
            call("hn_graph_insert", self._h, _ptr(batch.src), _ptr(batch.dst), batch.nE(), flags,
                 ctypes.byref(n), _stream())
        return n.value

    def erase(self, batch: BatchUpdate, removeBatchDuplicates: bool = False) -> int:
        flags = _lib.HN_DEVICE_PTRS | (_lib.HN_DEDUP_BATCH if removeBatchDuplicates else 0)
        n = ctypes.c_int64()
        call("hn_graph_erase", self._h, _ptr(batch.src), _ptr(batch.dst), batch.nE(), flags,
             ctypes.byref(n), _stream())
        return n.value

    def update_stats(self) -> dict:
        """What the last insert / erase did (hn_graph_update_stats): submitted
        and applied entries, touched and relocated vertices with their degree
        sums, and stored entries moved."""
        v = _lib.HnUpdateStats()
        call("hn_graph_update_stats", self._h, ctypes.byref(v), _stream())
        return {f: getattr(v, f) for f, _ in v._fields_ if f != "pad_"}

    # ------------------------------------------------------------------ cleanup
    def release(self):
        if self._h:
            _lib.load().hn_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class HornetStatic(Hornet):
    """hornet::gpu::HornetStatic<int> (hornet/include/Core/Static/HornetStatic.cuh:16-66):
    an immutable graph with the same queries and device view; insert and erase
    raise."""

    def insert(self, batch, removeBatchDuplicates: bool = False, removeGraphDuplicates: bool = False):
        raise HornetError(_lib.HN_E_INVALID_ARG, "HornetStatic.insert", "the graph is immutable")

    def erase(self, batch, removeBatchDuplicates: bool = False):
        raise HornetError(_lib.HN_E_INVALID_ARG, "HornetStatic.erase", "the graph is immutable")


class StaticAlgorithm:
    """primitives/HornetAlg.cuh:51-150: run/reset/release/validate."""

    def __init__(self, hornet: Hornet):
        self.hornet = hornet

    def run(self):
        raise NotImplementedError

    def reset(self):
        raise NotImplementedError

    def release(self):
        raise NotImplementedError

    def validate(self) -> bool:
        raise NotImplementedError


class BfsTopDown(StaticAlgorithm):
    """HornetsNest BfsTopDown (TopDown.cu:68-128).  ``validate`` needs a host
    reference, which the caller supplies (the reference's is a stub returning
    true, TopDown.cu:113-128)."""

    def __init__(self, hornet: Hornet):
        super().__init__(hornet)
        self.d_distances = torch.empty(max(hornet.nV(), 1), dtype=torch.int32, device=_dev())
        self.bfs_source = 0
        self.stats = _lib.HnBfsStats()

    def reset(self):
        self.d_distances.fill_(INF)

    def set_parameters(self, source: int):
        self.bfs_source = int(source)

    def run(self):
        call("hn_bfs_topdown", self.hornet.handle, self.bfs_source, _ptr(self.d_distances),
             ctypes.byref(self.stats), _stream())

    def release(self):
        self.d_distances = None

    def distances(self) -> torch.Tensor:
        return self.d_distances[: self.hornet.nV()]

    def validate(self, host_distances=None) -> bool:
        if host_distances is None:
            return True
        return bool(np.array_equal(self.distances().cpu().numpy(), np.asarray(host_distances)))


class TriangleCounting2(StaticAlgorithm):
    """HornetsNest TriangleCounting2 (triangle2.cu:24-175).  ``run`` counts on the
    graph as stored (the reference's driver stores the DIRECTED_BY_DEGREE
    orientation, TriangleTest2.cu:86-87); ``run_undirected`` orients the
    symmetric storage on the device first."""

    def __init__(self, hornet: Hornet):
        super().__init__(hornet)
        self.triPerVertex = None
        self.total = 0

    def init(self):
        self.triPerVertex = torch.zeros(max(self.hornet.nV(), 1), dtype=torch.int64, device=_dev())

    def reset(self):
        if self.triPerVertex is not None:
            self.triPerVertex.zero_()
        self.total = 0

    def run(self, work_factor: int = 1):
        if self.triPerVertex is None:
            self.init()
        t = ctypes.c_uint64()
        call("hn_tc2_stored", self.hornet.handle, ctypes.byref(t), _ptr(self.triPerVertex), _stream())
        self.total = t.value

    def run_undirected(self):
        if self.triPerVertex is None:
            self.init()
        t = ctypes.c_uint64()
        call("hn_tc_count", self.hornet.handle, ctypes.byref(t), _ptr(self.triPerVertex), _stream())
        self.total = t.value

    def countTriangles(self) -> int:
        return int(self.total)

    def copyTCToHost(self) -> np.ndarray:
        return self.triPerVertex[: self.hornet.nV()].cpu().numpy().view(np.uint64)

    def release(self):
        self.triPerVertex = None

    def validate(self) -> bool:
        return True


class BfsMulti:
    """Batched BFS of up to 64 sources per launch (bit-packed source masks,
    captured level loop).  ``run`` returns int32 distances of shape (k, nV)."""

    MAX_SOURCES = 64

    def __init__(self, hornet: Hornet):
        self.hornet = hornet
        self.stats = _lib.HnMsbfsStats()
        self._dist = None

    def _buf(self, k: int) -> torch.Tensor:
        n = k * max(self.hornet.nV(), 1)
        if self._dist is None or self._dist.numel() < n:
            self._dist = torch.empty(n, dtype=torch.int32, device=_dev())
        return self._dist[:n]

    def run(self, sources, depth_limits=None, out: torch.Tensor | None = None) -> torch.Tensor:
        src = np.ascontiguousarray(sources, dtype=np.int32)
        k = src.size
        if not 1 <= k <= self.MAX_SOURCES:
            raise HornetError(_lib.HN_E_INVALID_ARG, "BfsMulti.run", "1..64 sources per launch")
        dist = out if out is not None else self._buf(k)
        if depth_limits is None:
            call("hn_bfs_multi", self.hornet.handle, src.ctypes.data_as(ctypes.c_void_p), k, _ptr(dist),
                 ctypes.byref(self.stats), _stream())
        else:
            dl = np.ascontiguousarray(depth_limits, dtype=np.int32)
            call("hn_bfs_multi_depth", self.hornet.handle, src.ctypes.data_as(ctypes.c_void_p),
                 dl.ctypes.data_as(ctypes.c_void_p), k, _ptr(dist), ctypes.byref(self.stats), _stream())
        return dist.view(k, -1)[:, : self.hornet.nV()]

    def set_params(self, degree_threshold: int = 0, push_alpha: int = 0):
        call("hn_bfs_multi_params", self.hornet.handle, int(degree_threshold), int(push_alpha))

    def set_lb_tile(self, units_per_lane: int):
        """Expansion tile of BLOCK * units_per_lane work units (2, 4 or 6)."""
        call("hn_bfs_multi_lb_tile", self.hornet.handle, int(units_per_lane))

    def params(self):
        """(degree threshold, push alpha, units per lane) in use."""
        t, a, i = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        call("hn_bfs_multi_get_params", self.hornet.handle, ctypes.byref(t), ctypes.byref(a), ctypes.byref(i))
        return t.value, a.value, i.value

    def tune(self, sources):
        """Autotune the degree threshold, the push/pull switch and the LB tile;
        returns (threshold, alpha) (params() gives the tile as well)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        t, a = ctypes.c_int32(), ctypes.c_int32()
        call("hn_bfs_multi_tune", self.hornet.handle, src.ctypes.data_as(ctypes.c_void_p), src.size,
             _ptr(self._buf(src.size)), ctypes.byref(t), ctypes.byref(a), _stream())
        return t.value, a.value


def rmat_pairs_device(scale: int, seed: int, edge_factor: int = 16, first: int = 0, count=None):
    m = edge_factor << scale
    if count is None:
        count = m - first
    src = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    dst = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    call("hn_gen_rmat_pairs", int(scale), int(first), int(count), int(seed), _ptr(src), _ptr(dst),
         _stream())
    return src[:count], dst[:count]


class Server:
    """Continuous-batching request loop over a resident graph (hn_server_*).

    ``replay(trace)`` plays a list of requests at their arrival offsets and
    returns one result row per request (a numpy structured array)."""

    BFS, TWO_HOP, INSERT, ERASE = (_lib.HN_REQ_BFS, _lib.HN_REQ_TWO_HOP, _lib.HN_REQ_INSERT,
                                   _lib.HN_REQ_ERASE)

    def __init__(self, hornet: Hornet, max_batch: int = 64, slots: int = 2, mem_budget_bytes: int = 0,
                 plain: bool = False, min_window_queries: int = 0):
        self.hornet = hornet
        cfg = _lib.HnServeConfig(max_batch, slots, mem_budget_bytes, 1 if plain else 0, int(min_window_queries))
        self._s = ctypes.c_void_p()
        call("hn_server_create", hornet.handle, ctypes.byref(cfg), ctypes.byref(self._s))

    def slots(self) -> int:
        v = ctypes.c_int32()
        call("hn_server_slots", self._s, ctypes.byref(v))
        return v.value

    @staticmethod
    def slot_bytes(hornet: "Hornet", max_batch: int = 64) -> int:
        """Device bytes one launch slot takes from the memory budget."""
        v = ctypes.c_int64()
        call("hn_server_slot_bytes", hornet._h, int(max_batch), ctypes.byref(v))
        return v.value

    @staticmethod
    def _requests(trace):
        """trace: list of (kind, payload) with payload = source (queries) or
        (src, dst) int32 arrays (updates)."""
        keep = []
        reqs = (_lib.HnRequest * len(trace))()
        for i, (kind, payload) in enumerate(trace):
            r = reqs[i]
            r.kind = kind
            if kind in (_lib.HN_REQ_BFS, _lib.HN_REQ_TWO_HOP):
                r.source = int(payload)
            else:
                s = np.ascontiguousarray(payload[0], np.int32)
                d = np.ascontiguousarray(payload[1], np.int32)
                keep += [s, d]
                r.n = s.size
                r.src = s.ctypes.data
                r.dst = d.ctypes.data
        return reqs, keep

    def replay(self, trace, arrival_ns):
        reqs, keep = self._requests(trace)
        arr = np.ascontiguousarray(arrival_ns, np.int64)
        call("hn_server_replay", self._s, reqs, arr.ctypes.data_as(_lib.PI64), len(trace))
        del keep
        return self.results()

    def results(self) -> np.ndarray:
        n = ctypes.c_int64()
        call("hn_server_results", self._s, None, 0, ctypes.byref(n))
        buf = (_lib.HnResult * max(n.value, 1))()
        call("hn_server_results", self._s, buf, n.value, ctypes.byref(n))
        dt = np.dtype([(name, np.dtype(ct)) for name, ct in
                       [("id", np.int64), ("kind", np.int32), ("status", np.int32), ("t_arrive_ns", np.int64),
                        ("t_start_ns", np.int64), ("t_done_ns", np.int64), ("reached", np.int64),
                        ("dist_sum", np.int64), ("max_depth", np.int32), ("batch", np.int32),
                        ("applied", np.int64), ("version", np.uint64)]])
        return np.frombuffer(bytes(buf), dtype=dt, count=n.value).copy()

    def close(self):
        if self._s:
            _lib.load().hn_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chung_lu_pairs_device(n: int, count: int, seed: int, gamma: float = 2.1):
    src = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    dst = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    call("hn_gen_chung_lu_pairs", int(n), float(gamma), 0, int(count), int(seed), _ptr(src), _ptr(dst), _stream())
    return src[:count], dst[:count]


def uniform_pairs_device(n: int, count: int, seed: int):
    src = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    dst = torch.empty(max(count, 1), dtype=torch.int32, device=_dev())
    call("hn_gen_uniform_pairs", int(n), int(count), int(seed), _ptr(src), _ptr(dst), _stream())
    return src[:count], dst[:count]
