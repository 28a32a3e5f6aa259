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
* ``tc2_per_vertex``  -- sum equals host_tc on the degree-oriented graph; the
  credit rule (smaller-degree endpoint, ties to the destination, skips) by a
  hand-worked stored graph.
* ``directed_by_degree`` -- hand cases, including equal-degree ties (lower id
  keeps the edge); ``max_degree_id`` -- the hand-derived tie on G.mtx.
* the ``threads`` forms (``rmat_pairs``, ``chung_lu_pairs``, ``coo_to_csr``,
  ``host_tc`` / ``triangles``, ``bfs_many``) -- equal to the sequential
  functions on the same inputs (they only split the work; the triangle count
  intersects against a mark array instead of merging).
* ``insert``/``erase`` -- the reference's own identities
  (hornet/test/HornetInsertTest.cu:52-75: sorted(after) == sorted(before ++
  batch); hornet/test/HornetDeleteTest.cu:97-120: sorted(after ++ batch) ==
  before) plus set-union / set-difference brute force.
* ``rmat_pairs`` / ``uniform_pairs`` / ``chung_lu_pairs`` -- generators
  specified by this build (the reference has none): parity unpinned against the
  reference; pinned by distribution invariants and by the GPU generator.
* ``two_hop``         -- networkx ego-graph (radius 2) size.
* ``graph_std`` / ``read_market_std`` / ``read_dimacs9_std`` / ``read_binary``
  -- ``example/G.bin``, the reference's own binary dump of ``example/G.mtx``
  (default direction, file-order rows, binary layout); golden.json for the
  UNDIRECTED read; ``directed_by_degree`` for DIRECTED_BY_DEGREE; transpose
  and relabelling identities for ENABLE_INGOING, RM_SINGLETON and RANDOMIZE
  (``tests/test_graphio.py``).
* ``vertex_permutation`` -- specified by this build (the reference's RANDOMIZE
  is clock-seeded): parity unpinned against the reference; pinned as a
  permutation whose inverse restores the input graph.
* ``csr_meta`` / ``insert_meta`` / ``erase_meta`` -- a pure-Python list model
  of the same rules on random small multigraphs, the edge multisets of
  ``insert`` / ``erase`` (fields dropped), and hand cases for first-kept
  batch duplicates (``tests/test_oracle.py``).  Which stored copy of a
  multi-edge an erase removes is the library's choice (oldest first): parity
  unpinned against the reference, whose choice is unspecified.
"""
from __future__ import annotations

import collections
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
        L.orc_rmat_pairs_mt.restype = None
        L.orc_rmat_pairs_mt.argtypes = [c.c_int, c.c_int64, c.c_int64, c.c_uint64, c.c_int, i32p, i32p]
        L.orc_chung_lu_pairs_mt.restype = None
        L.orc_chung_lu_pairs_mt.argtypes = [c.c_int32, u64p, c.c_int64, c.c_int64, c.c_uint64, c.c_int,
                                            i32p, i32p]
        L.orc_coo_to_csr_mt.restype = c.c_int64
        L.orc_coo_to_csr_mt.argtypes = [c.c_int32, c.c_int64, i32p, i32p, c.c_int, c.c_int, c.c_int,
                                        i64p, i32p]
        L.orc_host_tc_mark_mt.restype = c.c_int64
        L.orc_host_tc_mark_mt.argtypes = [c.c_int32, i64p, i32p, c.c_int]
        _lib = L
    return _lib


def checker_threads() -> int:
    """Threads the parallel checker functions use by default: the CPUs this
    process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


# --------------------------------------------------------------------------
# generators
# --------------------------------------------------------------------------
def stream_key(seed: int) -> int:
    return int(lib().orc_stream_key(seed))


def rmat_pairs(scale: int, seed: int, edge_factor: int = 16, first: int = 0,
               count: int | None = None, threads: int = 1):
    m = edge_factor << scale
    if count is None:
        count = m - first
    src = np.empty(count, np.int32)
    dst = np.empty(count, np.int32)
    if threads > 1:
        lib().orc_rmat_pairs_mt(scale, first, count, seed, threads, src, dst)
    else:
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


def chung_lu_pairs(n: int, count: int, seed: int, gamma: float = 2.1, cum=None, threads: int = 1):
    if cum is None:
        cum = chung_lu_table(n, gamma)
    src = np.empty(count, np.int32)
    dst = np.empty(count, np.int32)
    if threads > 1:
        lib().orc_chung_lu_pairs_mt(n, cum, 0, count, seed, threads, src, dst)
    else:
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
               sort=True, threads: int = 1):
    """threads > 1 (sorted and deduplicated output only) splits the work over
    threads with the same result."""
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    m = src.size
    off = np.empty(n + 1, np.int64)
    adj = np.empty(max(2 * m, 1), np.int32)
    if threads > 1 and dedup and sort:
        k = lib().orc_coo_to_csr_mt(n, m, src, dst, int(symmetrize), int(drop_self_loops), threads, off, adj)
        if k < 0:
            raise MemoryError("orc_coo_to_csr_mt: out of host memory")
    else:
        k = lib().orc_coo_to_csr(n, m, src, dst, int(symmetrize), int(drop_self_loops),
                                 int(dedup), int(sort), off, adj)
    if k == adj.size:
        return off, adj
    return off, adj[:k].copy()


def rmat_graph(scale: int, seed: int, edge_factor: int = 16, threads: int = 1):
    """R-MAT(a=.57,b=.19,c=.19), self loops dropped, symmetrized, deduped, sorted."""
    s, d = rmat_pairs(scale, seed, edge_factor, threads=threads)
    return coo_to_csr(1 << scale, s, d, threads=threads)


def chung_lu_graph(n: int, avg_degree: int, gamma: float, seed: int, threads: int = 1):
    """Chung-Lu graph as the device builds it: n * avg_degree / 2 pairs, self
    loops dropped, symmetrized, deduped, sorted."""
    s, d = chung_lu_pairs(n, n * avg_degree // 2, seed, gamma, threads=threads)
    return coo_to_csr(n, s, d, threads=threads)


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


def bfs_many(off, adj, sources, threads: int | None = None):
    """orc_bfs from each source, the sources spread over threads (the C code
    runs without the interpreter lock); returns [(dist, edges_scanned)]."""
    from concurrent.futures import ThreadPoolExecutor
    off = np.ascontiguousarray(off, np.int64)
    adj = np.ascontiguousarray(adj, np.int32)
    srcs = [int(s) for s in sources]
    with ThreadPoolExecutor(max_workers=threads or checker_threads()) as ex:
        return list(ex.map(lambda s: bfs(off, adj, s), srcs))


def host_tc(off, adj, threads: int = 1) -> int:
    n = off.size - 1
    off = np.ascontiguousarray(off, np.int64)
    adj = np.ascontiguousarray(adj, np.int32)
    if threads > 1:
        r = int(lib().orc_host_tc_mark_mt(n, off, adj, threads))
        if r < 0:
            raise MemoryError("orc_host_tc_mark_mt: out of host memory")
        return r
    return int(lib().orc_host_tc(n, off, adj))


def triangles(off, adj, threads: int = 1) -> int:
    """Triangle count of the undirected simple graph stored symmetrically."""
    o, a = directed_by_degree(off, adj)
    return host_tc(o, a, threads)


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


# --------------------------------------------------------------------------
# GraphStd: file direction rules, COO -> CSR and the binary graph format
# --------------------------------------------------------------------------
# StructureProp values (xlib/include/Graph/GraphBase.hpp:92-93)
GS_DIRECTED, GS_UNDIRECTED, GS_ENABLE_INGOING = 1, 2, 4
# ParsingProp values (GraphBase.hpp:46-48)
P_RANDOMIZE, P_SORT, P_PRINT_INFO, P_RM_SINGLETON, P_DIRECTED_BY_DEGREE = 1, 2, 4, 8, 16


def vertex_permutation(n: int, seed: int) -> np.ndarray:
    """RANDOMIZE relabelling as the library specifies it (the reference shuffles
    with a clock seed, GraphStd.cpp:217-229, so no fixed permutation of its can
    be matched): Fisher-Yates from i = n-1 down to 1, j = uniform_below(draw(i),
    i+1) of SplitMix64 stream `seed`; vertex v becomes perm[v]."""
    key = stream_key(seed)
    perm = np.arange(n, dtype=np.int64)
    for i in range(n - 1, 0, -1):
        r = int(lib().orc_draw(key, i))
        j = ((r >> 32) * (i + 1)) >> 32
        perm[i], perm[j] = perm[j], perm[i]
    return perm


def _stable_csr(n, a, b):
    """Rows a, entries b, each row in input order (GraphStd.cpp:333-341)."""
    order = np.argsort(a, kind="stable")
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, a + 1, 1)
    return np.cumsum(off), b[order].astype(np.int32)


def graph_std(n, src, dst, file_direction, stored_undirected, structure=0, parse=0, seed=0):
    """GraphStd::read after the file's entries are in COO form: direction
    resolution (GraphStd.cpp:85-95), mirror of stored-undirected /
    directed-to-undirected input (:175-190, self loops not mirrored), sort +
    unique for directed-to-undirected (:192-206), RANDOMIZE (:217-229, with
    vertex_permutation), degree count (:233-245), DIRECTED_BY_DEGREE (:247-304,
    deg(u) < deg(v) or equal degrees and u < v), SORT (:306-310),
    RM_SINGLETON (:312-331), CSR placement in COO order (:333-351).

    Where the reference's behaviour is undefined this restatement fixes it:
    RM_SINGLETON drops the edges that touch a removed vertex (the reference
    leaves their labels unset), and an UNDIRECTED graph's in-CSR is its out-CSR
    (the reference zeroes its in-degrees after DIRECTED_BY_DEGREE).  A file
    stored undirected read into a DIRECTED structure (random edge removal,
    :96-101) is rejected.  Returns a dict n, structure, off, adj, in_off, in_adj
    (in_* None unless DIRECTED with ENABLE_INGOING)."""
    s = np.asarray(src, np.int64).copy()
    d = np.asarray(dst, np.int64).copy()
    if not structure & (GS_DIRECTED | GS_UNDIRECTED):
        structure |= file_direction
    if file_direction == GS_UNDIRECTED and structure & GS_DIRECTED:
        raise ValueError("undirected file into a directed structure")
    d2u = file_direction == GS_DIRECTED and bool(structure & GS_UNDIRECTED)
    if d2u or stored_undirected:
        keep = s != d
        s, d = np.concatenate([s, d[keep]]), np.concatenate([d, s[keep]])
    if d2u:
        pairs = sorted(set(zip(s.tolist(), d.tolist())))
        s = np.array([p[0] for p in pairs], np.int64)
        d = np.array([p[1] for p in pairs], np.int64)
    if parse & P_RANDOMIZE:
        perm = vertex_permutation(n, seed)
        s, d = perm[s], perm[d]
    directed = bool(structure & GS_DIRECTED)
    ingoing = directed and bool(structure & GS_ENABLE_INGOING)
    reverse = bool(structure & (GS_ENABLE_INGOING | GS_UNDIRECTED))

    def degrees():
        o = np.bincount(s, minlength=n)
        return o, (np.bincount(d, minlength=n) if ingoing else o)

    outd, ind = degrees()
    if parse & P_DIRECTED_BY_DEGREE:
        keep = (outd[s] < outd[d]) | ((outd[s] == outd[d]) & (s < d))
        s, d = s[keep], d[keep]
        outd, ind = degrees()
    if parse & P_SORT and (not d2u or parse & P_RANDOMIZE):
        order = np.lexsort((d, s))
        s, d = s[order], d[order]
    if parse & P_RM_SINGLETON and reverse:
        kept = (outd > 0) & (ind > 0)
        label = np.cumsum(kept) - 1
        both = kept[s] & kept[d]
        s, d = label[s[both]], label[d[both]]
        n = int(kept.sum())
    off, adj = _stable_csr(n, s, d)
    in_off = in_adj = None
    if ingoing:
        in_off, in_adj = _stable_csr(n, d, s)
    return {"n": n, "structure": structure, "off": off, "adj": adj, "in_off": in_off,
            "in_adj": in_adj}


def read_market_std(path: str, structure=0, parse=0, seed=0):
    """GraphStd::read of a MatrixMarket file: header 'symmetric' -> UNDIRECTED
    stored undirected, else DIRECTED (GraphBase.cpp:214-238)."""
    n, s, d, sym = read_market(path)
    return graph_std(n, s, d, GS_UNDIRECTED if sym else GS_DIRECTED, sym, structure, parse, seed)


def read_dimacs9_std(path: str, structure=0, parse=0, seed=0):
    """GraphStd::read of a DIMACS9 file: UNDIRECTED, arcs as listed, nothing
    mirrored (GraphBase.cpp:252-262)."""
    n, s, d = read_dimacs9(path)
    return graph_std(n, s, d, GS_UNDIRECTED, False, structure, parse, seed)


def read_binary(path: str):
    """Binary graph (GraphStd.cpp:360-421 writer, GraphStdRead.cpp:206-236
    reader): the class id type_name<vid_t>() + type_name<eoff_t>() ("intint"
    or "intlong"), nV, nE, the 16-byte StructureProp object (its uint64 state at
    byte 8), then out offsets [nV+1], (in offsets), out edges [nE], (in edges)
    -- the in arrays present when the state is DIRECTED | ENABLE_INGOING.
    Degrees are offsets[i+1] - offsets[i] (the reference reader subtracts
    offsets[i-1])."""
    raw = open(path, "rb").read()
    for cid, ot in ((b"intint", np.int32), (b"intlong", np.int64)):
        if raw.startswith(cid):
            break
    else:
        raise ValueError("unknown class identifier")
    p = len(cid)
    nv = int(np.frombuffer(raw, np.int32, 1, p)[0]); p += 4
    ne = int(np.frombuffer(raw, ot, 1, p)[0]); p += np.dtype(ot).itemsize
    state = int(np.frombuffer(raw, np.uint64, 1, p + 8)[0]); p += 16
    twice = (state & GS_DIRECTED) and (state & GS_ENABLE_INGOING)
    off = np.frombuffer(raw, ot, nv + 1, p).astype(np.int64); p += (nv + 1) * np.dtype(ot).itemsize
    in_off = in_adj = None
    if twice:
        in_off = np.frombuffer(raw, ot, nv + 1, p).astype(np.int64); p += (nv + 1) * np.dtype(ot).itemsize
    adj = np.frombuffer(raw, np.int32, ne, p).copy(); p += 4 * ne
    if twice:
        in_adj = np.frombuffer(raw, np.int32, ne, p).copy(); p += 4 * ne
    if p != len(raw):
        raise ValueError("file size does not match its header")
    return {"n": nv, "structure": state, "off": off, "adj": adj, "in_off": in_off, "in_adj": in_adj}


# --------------------------------------------------------------------------
# edge meta-data: rows sorted by destination, equal destinations in
# insertion order, fields carried with their edges
# --------------------------------------------------------------------------
def _meta_rows(n, s, d, order, fields):
    """CSR of entries (s, d) ordered by (s, d, order), fields permuted along."""
    p = np.lexsort((order, d, s))
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, np.asarray(s, np.int64) + 1, 1)
    return np.cumsum(off), np.asarray(d, np.int32)[p], [np.asarray(f)[p] for f in fields]


def csr_meta(off, adj, fields):
    """Hornet(HornetInit) with insertEdgeData (HornetInit.cuh:75-97): each row
    sorted by destination, equal destinations in input order."""
    n = off.size - 1
    s = np.repeat(np.arange(n), np.diff(off))
    return _meta_rows(n, s, adj, np.arange(adj.size), fields)


def insert_meta(off, adj, fields, bs, bd, bfields, dedup_batch=False, dedup_graph=False):
    """Hornet::insert of a batch with meta-data: the batch is stably sorted by
    (src, dst) with its fields (SoAPtr.i.cuh:581-613); removeBatchDuplicates
    keeps the first of equal pairs (unique_copy, BatchUpdate.i.cuh:250-335);
    removeGraphDuplicates drops the first batch copy of each pair already
    stored (as orc_insert, after BatchUpdateKernels.cuh:27-58; the stored
    fields stay).  New copies of a destination follow the stored ones.
    Returns (off, adj, fields, added)."""
    n = off.size - 1
    bs = np.asarray(bs, np.int64)
    bd = np.asarray(bd, np.int64)
    keep = np.ones(bs.size, bool)
    p = np.lexsort((np.arange(bs.size), bd, bs))
    first = np.ones(bs.size, bool)
    first[1:] = (bs[p][1:] != bs[p][:-1]) | (bd[p][1:] != bd[p][:-1])
    if dedup_batch:
        keep[:] = False
        keep[p[first]] = True
    if dedup_graph:
        s0 = np.repeat(np.arange(n), np.diff(off)).astype(np.int64)
        stored = set(zip(s0.tolist(), np.asarray(adj).tolist()))
        for i in p[first]:
            if (int(bs[i]), int(bd[i])) in stored:
                keep[i] = False
    s0 = np.repeat(np.arange(n), np.diff(off)).astype(np.int64)
    s = np.concatenate([s0, bs[keep]])
    d = np.concatenate([np.asarray(adj, np.int64), bd[keep]])
    order = np.concatenate([np.arange(adj.size), adj.size + np.flatnonzero(keep)])
    fl = [np.concatenate([np.asarray(f), np.asarray(bf)[keep]]) for f, bf in zip(fields, bfields)]
    o, a, f = _meta_rows(n, s, d, order, fl)
    return o, a, f, int(keep.sum())


def erase_meta(off, adj, fields, bs, bd, dedup_batch=False):
    """Hornet::erase on a graph with meta-data: a distinct (u, v) of batch
    multiplicity c removes the min(c, stored) oldest stored copies (the
    library's choice; the reference's copy choice is unspecified); fields of
    the survivors stay with them.  Returns (off, adj, fields, removed)."""
    n = off.size - 1
    cnt = collections.Counter(zip(np.asarray(bs).tolist(), np.asarray(bd).tolist()))
    if dedup_batch:
        cnt = collections.Counter({k: 1 for k in cnt})
    s0 = np.repeat(np.arange(n), np.diff(off))
    keep = np.ones(adj.size, bool)
    for j in range(adj.size):          # rows hold equal destinations oldest first
        k = (int(s0[j]), int(adj[j]))
        if cnt.get(k, 0) > 0:
            cnt[k] -= 1
            keep[j] = False
    o = np.zeros(n + 1, np.int64)
    np.add.at(o, s0[keep] + 1, 1)
    return (np.cumsum(o), np.asarray(adj)[keep], [np.asarray(f)[keep] for f in fields],
            int((~keep).sum()))
