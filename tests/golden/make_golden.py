"""Regenerate tests/golden/golden.json from the reference's example graphs.

Inputs are the data files the reference ships in example/ (copied here as
fixtures): G.mtx, G_cc2.mtx (MatrixMarket, general) and rome99.gr (DIMACS9).
They are read the way the reference drivers read them (GraphStd with
structure UNDIRECTED for the .mtx files -> symmetrize + sort + unique,
xlib/src/Graph/GraphStd.cpp:173-206; DIMACS9 arcs as stored, then sorted).

Expected outputs come from the oracle and are cross-checked here against
networkx (an independent implementation) before being written:
  * BFS distances from the max-out-degree vertex (BFSTest.cu:28) and from
    vertex 0, INT32_MAX for unreached;
  * the triangle count of the undirected simple graph, and TriangleCounting2's
    total on the DIRECTED_BY_DEGREE orientation (TriangleTest2.cu:86-87).
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

INF = 2**31 - 1


def nx_dist(off, adj, src):
    n = off.size - 1
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    s = np.repeat(np.arange(n), np.diff(off))
    G.add_edges_from(zip(s.tolist(), adj.tolist()))
    d = nx.single_source_shortest_path_length(G, src)
    out = np.full(n, INF, np.int64)
    for k, v in d.items():
        out[k] = v
    return out


def nx_triangles(off, adj):
    n = off.size - 1
    G = nx.Graph()
    G.add_nodes_from(range(n))
    s = np.repeat(np.arange(n), np.diff(off))
    G.add_edges_from((a, b) for a, b in zip(s.tolist(), adj.tolist()) if a != b)
    return sum(nx.triangles(G).values()) // 3


def load(name):
    path = os.path.join(HERE, name)
    if name.endswith(".mtx"):
        n, src, dst, _ = O.read_market(path)
        off, adj = O.coo_to_csr(n, src, dst, symmetrize=True, drop_self_loops=False, dedup=True)
    else:
        n, src, dst = O.read_dimacs9(path)
        off, adj = O.coo_to_csr(n, src, dst, symmetrize=False, drop_self_loops=False, dedup=False,
                                sort=True)
    return off, adj


def main():
    out = {}
    for name in ("G.mtx", "G_cc2.mtx", "rome99.gr"):
        off, adj = load(name)
        root = O.max_degree_id(off)
        entry = {"nv": int(off.size - 1), "ne": int(off[-1]), "max_degree_id": root,
                 "offsets_sha": None}
        for tag, src in (("bfs_maxdeg", root), ("bfs_0", 0)):
            d, _ = O.bfs(off, adj, src)
            ref = nx_dist(off, adj, src)
            assert np.array_equal(d.astype(np.int64), ref), (name, tag)
            entry[tag] = {"source": int(src), "dist": d.tolist()}
        # symmetric view for the triangle count (rome99 arcs are symmetrised here)
        s = np.repeat(np.arange(off.size - 1), np.diff(off)).astype(np.int32)
        so, sa = O.coo_to_csr(off.size - 1, s, adj, symmetrize=True, drop_self_loops=True, dedup=True)
        tri = O.triangles(so, sa)
        assert tri == nx_triangles(so, sa), name
        do, da = O.directed_by_degree(so, sa)
        tc2 = int(O.tc2_per_vertex(do, da).sum())
        assert tc2 == tri, name
        entry["triangles"] = int(tri)
        entry["sym_ne"] = int(so[-1])
        entry["offsets"] = off.tolist() if off.size <= 64 else None
        entry["adj"] = adj.tolist() if adj.size <= 64 else None
        entry.pop("offsets_sha")
        out[name] = entry
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print({k: (v["nv"], v["ne"], v["triangles"]) for k, v in out.items()})


if __name__ == "__main__":
    main()
