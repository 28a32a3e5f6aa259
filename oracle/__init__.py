"""CPU oracle for the Hornet hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP product in ``hornet_amd``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only to compute expected results or a timed
host baseline -- never as the thing measured or shipped.

The numerics live in ``oracle/oracle.c`` (built by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``); each C function cites the reference file:line
it restates.  This module is a thin numpy/ctypes wrapper.

Pinning status (what independently checks each function, see
``tests/test_oracle.py``):

* ``coo_to_csr``      -- golden fixtures from ``example/G.mtx`` and
  ``example/rome99.gr``; networkx graph construction.
* ``bfs``             -- networkx single-source shortest path lengths; BFS
  invariants (|d(u)-d(v)| <= 1 across edges, parent at d-1).
* ``host_tc``         -- networkx triangle count; dense trace(A^3)/6.
* ``tc2_per_vertex``  -- sum equals host_tc on the degree-oriented graph.
* ``insert``/``erase`` -- the reference's own identities
  (hornet/test/HornetInsertTest.cu:52-75: sorted(after) == sorted(before ++
  batch); hornet/test/HornetDeleteTest.cu:97-120: sorted(after ++ batch) ==
  before) plus set-union / set-difference brute force.
* ``rmat_pairs`` / ``uniform_pairs`` / ``chung_lu_pairs`` -- generators
  specified by this build (the reference has none): parity unpinned against the
  reference; pinned by distribution invariants and by the GPU generator.
* ``two_hop``         -- networkx ego-graph (radius 2) size.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        c = ctypes
        L.orc_stream_key.restype = c.c_uint64
        L.orc_stream_key.argtypes = [c.c_uint64]
        L.orc_draw.restype = c.c_uint64
        L.orc_draw.argtypes = [c.c_uint64, c.c_uint64]
        L.orc_rmat_pairs.restype = None
        L.orc_rmat_pairs.argtypes = [c.c_int, c.c_int64, c.c_int64, c.c_uint64, i32p, i32p]
        L.orc_uniform_pairs.restype = None
        L.orc_uniform_pairs.argtypes = [c.c_int32, c.c_int64, c.c_uint64, i32p, i32p]
        L.orc_pick_sources.restype = c.c_int64
        L.orc_pick_sources.argtypes = [c.c_int32, i64p, c.c_int32, c.c_uint64, i32p]
        L.orc_coo_to_csr.restype = c.c_int64
        L.orc_coo_to_csr.argtypes = [c.c_int32, c.c_int64, i32p, i32p, c.c_int, c.c_int,
                                     c.c_int, c.c_int, i64p, i32p]
        L.orc_directed_by_degree.restype = c.c_int64
        L.orc_directed_by_degree.argtypes = [c.c_int32, i64p, i32p, i64p, i32p]
        L.orc_bfs.restype = c.c_int64
        L.orc_bfs.argtypes = [c.c_int32, i64p, i32p, c.c_int32, i32p,
                              ctypes.POINTER(c.c_int64)]
        L.orc_host_tc.restype = c.c_int64
        L.orc_host_tc.argtypes = [c.c_int32, i64p, i32p]
        L.orc_tc2_per_vertex.restype = None
        L.orc_tc2_per_vertex.argtypes = [c.c_int32, i64p, i32p, u64p]
        L.orc_max_degree_id.restype = c.c_int32
        L.orc_max_degree_id.argtypes = [c.c_int32, i64p]
        L.orc_insert.restype = c.c_int64
        L.orc_insert.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int64, c.c_int,
                                 c.c_int, i64p, i32p, ctypes.POINTER(c.c_int64)]
        L.orc_erase.restype = c.c_int64
        L.orc_erase.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int64, c.c_int,
                                i64p, i32p, ctypes.POINTER(c.c_int64)]
        L.orc_two_hop.restype = c.c_int64
        L.orc_two_hop.argtypes = [c.c_int32, i64p, i32p, c.c_int32, u8p]
        L.orc_chung_lu_table.restype = None
        L.orc_chung_lu_table.argtypes = [c.c_int32, c.c_double, u64p]
        L.orc_chung_lu_pairs.restype = None
        L.orc_chung_lu_pairs.argtypes = [c.c_int32, u64p, c.c_int64, c.c_int64, c.c_uint64,
                                         i32p, i32p]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# generators
# --------------------------------------------------------------------------
def stream_key(seed: int) -> int:
    return int(lib().orc_stream_key(seed))


def rmat_pairs(scale: int, seed: int, edge_factor: int = 16, first: int = 0,
               count: int | None = None):
    m = edge_factor << scale
    if count is None:
        count = m - first
    src = np.empty(count, np.int32)
    dst = np.empty(count, np.int32)
    lib().orc_rmat_pairs(scale, first, count, seed, src, dst)
    return src, dst


def uniform_pairs(n: int, count: int, seed: int):
    src = np.empty(count, np.int32)
    dst = np.empty(count, np.int32)
    lib().orc_uniform_pairs(n, count, seed, src, dst)
    return src, dst


def chung_lu_table(n: int, gamma: float = 2.1):
    cum = np.empty(n + 1, np.uint64)
    lib().orc_chung_lu_table(n, gamma, cum)
    return cum


def chung_lu_pairs(n: int, count: int, seed: int, gamma: float = 2.1, cum=None):
    if cum is None:
        cum = chung_lu_table(n, gamma)
    src = np.empty(count, np.int32)
    dst = np.empty(count, np.int32)
    lib().orc_chung_lu_pairs(n, cum, 0, count, seed, src, dst)
    return src, dst


def pick_sources(off: np.ndarray, k: int, seed: int) -> np.ndarray:
    n = off.size - 1
    out = np.empty(k, np.int32)
    r = lib().orc_pick_sources(n, np.ascontiguousarray(off, np.int64), k, seed, out)
    if r < 0:
        raise ValueError("graph has no non-isolated vertex")
    return out


# --------------------------------------------------------------------------
# graph construction
# --------------------------------------------------------------------------
def coo_to_csr(n: int, src, dst, symmetrize=True, drop_self_loops=True, dedup=True,
               sort=True):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    m = src.size
    off = np.empty(n + 1, np.int64)
    adj = np.empty(max(2 * m, 1), np.int32)
    k = lib().orc_coo_to_csr(n, m, src, dst, int(symmetrize), int(drop_self_loops),
                             int(dedup), int(sort), off, adj)
    return off, adj[:k].copy()


def rmat_graph(scale: int, seed: int, edge_factor: int = 16):
    """R-MAT(a=.57,b=.19,c=.19), self loops dropped, symmetrized, deduped, sorted."""
    s, d = rmat_pairs(scale, seed, edge_factor)
    return coo_to_csr(1 << scale, s, d)


def directed_by_degree(off, adj):
    n = off.size - 1
    o = np.empty(n + 1, np.int64)
    a = np.empty(max(adj.size, 1), np.int32)
    k = lib().orc_directed_by_degree(n, np.ascontiguousarray(off, np.int64),
                                     np.ascontiguousarray(adj, np.int32), o, a)
    return o, a[:k].copy()


# --------------------------------------------------------------------------
# algorithms
# --------------------------------------------------------------------------
def bfs(off, adj, source: int):
    n = off.size - 1
    dist = np.empty(n, np.int32)
    scanned = ctypes.c_int64(0)
    lib().orc_bfs(n, np.ascontiguousarray(off, np.int64), np.ascontiguousarray(adj, np.int32),
                  int(source), dist, ctypes.byref(scanned))
    return dist, int(scanned.value)


def host_tc(off, adj) -> int:
    n = off.size - 1
    return int(lib().orc_host_tc(n, np.ascontiguousarray(off, np.int64),
                                 np.ascontiguousarray(adj, np.int32)))


def triangles(off, adj) -> int:
    """Triangle count of the undirected simple graph stored symmetrically."""
    o, a = directed_by_degree(off, adj)
    return host_tc(o, a)


def tc2_per_vertex(off, adj):
    n = off.size - 1
    out = np.empty(n, np.uint64)
    lib().orc_tc2_per_vertex(n, np.ascontiguousarray(off, np.int64),
                             np.ascontiguousarray(adj, np.int32), out)
    return out


def max_degree_id(off) -> int:
    return int(lib().orc_max_degree_id(off.size - 1, np.ascontiguousarray(off, np.int64)))


def insert(off, adj, bsrc, bdst, dedup_batch=False, dedup_graph=False):
    n = off.size - 1
    b = len(bsrc)
    o = np.empty(n + 1, np.int64)
    a = np.empty(adj.size + b + 1, np.int32)
    added = ctypes.c_int64(0)
    k = lib().orc_insert(n, np.ascontiguousarray(off, np.int64), np.ascontiguousarray(adj, np.int32),
                         np.ascontiguousarray(bsrc, np.int32), np.ascontiguousarray(bdst, np.int32),
                         b, int(dedup_batch), int(dedup_graph), o, a, ctypes.byref(added))
    return o, a[:k].copy(), int(added.value)


def erase(off, adj, bsrc, bdst, dedup_batch=False):
    n = off.size - 1
    b = len(bsrc)
    o = np.empty(n + 1, np.int64)
    a = np.empty(adj.size + 1, np.int32)
    removed = ctypes.c_int64(0)
    k = lib().orc_erase(n, np.ascontiguousarray(off, np.int64), np.ascontiguousarray(adj, np.int32),
                        np.ascontiguousarray(bsrc, np.int32), np.ascontiguousarray(bdst, np.int32),
                        b, int(dedup_batch), o, a, ctypes.byref(removed))
    return o, a[:k].copy(), int(removed.value)


def two_hop(off, adj, v: int) -> int:
    n = off.size - 1
    mark = np.zeros(n, np.uint8)
    return int(lib().orc_two_hop(n, np.ascontiguousarray(off, np.int64),
                                 np.ascontiguousarray(adj, np.int32), int(v), mark))


# --------------------------------------------------------------------------
# reference file readers (restated for the fixtures)
# --------------------------------------------------------------------------
def read_market(path: str):
    """MatrixMarket coordinate reader (xlib/src/Graph/GraphStdRead.cpp:45-61,
    header GraphBase.cpp:214-238): 1-based ids; returns (n, src, dst, symmetric)."""
    with open(path) as f:
        header = f.readline()
        symmetric = "symmetric" in header
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        rows, cols, nl = (int(x) for x in line.split()[:3])
        src = np.empty(nl, np.int32)
        dst = np.empty(nl, np.int32)
        for i in range(nl):
            p = f.readline().split()
            src[i] = int(p[0]) - 1
            dst[i] = int(p[1]) - 1
    return max(rows, cols), src, dst, symmetric


def read_dimacs9(path: str):
    """DIMACS9 '.gr' reader (GraphStdRead.cpp:87-105, header GraphBase.cpp:252-262)."""
    src, dst = [], []
    n = 0
    with open(path) as f:
        for line in f:
            if line.startswith("p"):
                n = int(line.split()[2])
            elif line.startswith("a"):
                p = line.split()
                src.append(int(p[1]) - 1)
                dst.append(int(p[2]) - 1)
    return n, np.array(src, np.int32), np.array(dst, np.int32)
