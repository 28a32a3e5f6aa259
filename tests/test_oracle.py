"""Pins for the CPU oracle (no GPU).

Each oracle function is checked against something that does not restate it:
the golden fixtures built from the reference's example graphs, networkx, dense
linear algebra, and the identities the reference's own update tests assert
(hornet/test/HornetInsertTest.cu:52-75, hornet/test/HornetDeleteTest.cu:97-120).
"""
import collections
import json
import os

import networkx as nx
import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
INF = 2**31 - 1


def to_nx(off, adj, directed=False):
    n = off.size - 1
    G = nx.DiGraph() if directed else nx.Graph()
    G.add_nodes_from(range(n))
    s = np.repeat(np.arange(n), np.diff(off))
    G.add_edges_from(zip(s.tolist(), adj.tolist()))
    return G


def coo(off, adj):
    s = np.repeat(np.arange(off.size - 1), np.diff(off)).astype(np.int32)
    return s, adj.astype(np.int32)


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def rmat10():
    return O.rmat_graph(10, seed=7)


# ---------------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", ["G.mtx", "G_cc2.mtx", "rome99.gr"])
def test_fixture_bfs_and_triangles(golden, name):
    g = golden[name]
    path = os.path.join(GOLD, name)
    if name.endswith(".mtx"):
        n, s, d, _ = O.read_market(path)
        off, adj = O.coo_to_csr(n, s, d, symmetrize=True, drop_self_loops=False, dedup=True)
    else:
        n, s, d = O.read_dimacs9(path)
        off, adj = O.coo_to_csr(n, s, d, symmetrize=False, drop_self_loops=False, dedup=False, sort=True)
    assert off.size - 1 == g["nv"] and off[-1] == g["ne"]
    assert O.max_degree_id(off) == g["max_degree_id"]
    for tag in ("bfs_maxdeg", "bfs_0"):
        dist, _ = O.bfs(off, adj, g[tag]["source"])
        assert dist.tolist() == g[tag]["dist"]
        # independent check: networkx shortest path lengths on the stored arcs
        ref = nx.single_source_shortest_path_length(to_nx(off, adj, directed=True), g[tag]["source"])
        exp = np.full(off.size - 1, INF)
        for k, v in ref.items():
            exp[k] = v
        assert np.array_equal(dist, exp)
    so, sa = O.coo_to_csr(off.size - 1, *coo(off, adj), symmetrize=True, drop_self_loops=True)
    assert O.triangles(so, sa) == g["triangles"]


def test_gmtx_csr_matches_reader_semantics(golden):
    # G.mtx has 12 arcs, one reciprocal pair (3,4)/(4,3): 24 - 2 = 22 after sort+unique
    g = golden["G.mtx"]
    assert g["ne"] == 22
    off, adj = np.array(g["offsets"]), np.array(g["adj"])
    for v in range(off.size - 1):
        seg = adj[off[v]:off[v + 1]]
        assert np.all(np.diff(seg) > 0)


# ---------------------------------------------------------------- BFS
def test_bfs_vs_networkx_rmat(rmat10):
    off, adj = rmat10
    G = to_nx(off, adj)
    for src in [O.max_degree_id(off), 3, 100, 1023]:
        dist, scanned = O.bfs(off, adj, src)
        ref = nx.single_source_shortest_path_length(G, src)
        exp = np.full(off.size - 1, INF)
        for k, v in ref.items():
            exp[k] = v
        assert np.array_equal(dist, exp)
        deg = np.diff(off)
        assert scanned == int(deg[dist != INF].sum())


def test_bfs_invariants(rmat10):
    off, adj = rmat10
    src = O.max_degree_id(off)
    dist, _ = O.bfs(off, adj, src)
    s, d = coo(off, adj)
    ds, dd = dist[s].astype(np.int64), dist[d].astype(np.int64)
    reached = (ds != INF) & (dd != INF)
    assert np.all(np.abs(ds[reached] - dd[reached]) <= 1)
    # reached and unreached never share an edge (symmetric graph)
    assert np.all((ds == INF) == (dd == INF))
    # every reached non-source vertex has a neighbour one level closer
    has_parent = np.zeros(off.size - 1, bool)
    ok = reached & (dd == ds + 1)
    has_parent[d[ok]] = True
    rv = np.flatnonzero(dist != INF)
    assert np.all(has_parent[rv[rv != src]])
    assert dist[src] == 0


# ---------------------------------------------------------------- triangles
def test_triangles_vs_networkx_and_dense(rmat10):
    off, adj = rmat10
    tri = O.triangles(off, adj)
    assert tri == sum(nx.triangles(to_nx(off, adj)).values()) // 3
    n = off.size - 1
    A = np.zeros((n, n), np.float64)
    s, d = coo(off, adj)
    A[s, d] = 1
    assert tri == int(round(np.trace(A @ A @ A) / 6))


def test_tc2_semantics_on_oriented_graph(rmat10):
    off, adj = rmat10
    do, da = O.directed_by_degree(off, adj)
    # each undirected edge oriented exactly once
    assert da.size * 2 == adj.size
    per = O.tc2_per_vertex(do, da)
    assert int(per.sum()) == O.host_tc(do, da) == O.triangles(off, adj)
    # on the symmetric storage every triangle is seen from all six arcs, and
    # the smaller-degree endpoint credit rule still sums over all arcs
    assert int(O.tc2_per_vertex(off, adj).sum()) == 6 * O.triangles(off, adj)


def test_directed_by_degree_rule():
    # path 0-1-2 plus 1-3: deg(1)=3 is the largest, so every edge points to 1
    off = np.array([0, 1, 4, 5, 6], np.int64)
    adj = np.array([1, 0, 2, 3, 1, 1], np.int32)
    do, da = O.directed_by_degree(off, adj)
    assert do.tolist() == [0, 1, 1, 2, 3] and da.tolist() == [1, 1, 1]


# ---------------------------------------------------------------- updates
def multiset(off, adj):
    return collections.Counter(zip(*[x.tolist() for x in coo(off, adj)]))


def test_insert_identity_reference(rmat10):
    # HornetInsertTest.cu:52-75: without dedup, sorted(after) == sorted(before ++ batch)
    off, adj = rmat10
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 500, seed=11)
    o2, a2, added = O.insert(off, adj, bs, bd)
    assert added == 500
    s0, d0 = coo(off, adj)
    before = np.concatenate([(s0.astype(np.int64) << 32) | d0, (bs.astype(np.int64) << 32) | bd])
    s1, d1 = coo(o2, a2)
    after = (s1.astype(np.int64) << 32) | d1
    assert np.array_equal(np.sort(before), after)  # after is already sorted


def test_erase_identity_reference(rmat10):
    # HornetDeleteTest.cu:97-120: erase distinct stored edges, then
    # sorted(after ++ batch) == before
    off, adj = rmat10
    s0, d0 = coo(off, adj)
    rng = np.random.default_rng(5)
    pick = rng.choice(adj.size, 300, replace=False)
    o2, a2, removed = O.erase(off, adj, s0[pick], d0[pick])
    assert removed == 300
    s1, d1 = coo(o2, a2)
    merged = np.concatenate([(s1.astype(np.int64) << 32) | d1,
                             (s0[pick].astype(np.int64) << 32) | d0[pick]])
    assert np.array_equal(np.sort(merged), (s0.astype(np.int64) << 32) | d0)


def test_insert_dedup_flags_set_semantics(rmat10):
    off, adj = rmat10
    n = off.size - 1
    s0, d0 = coo(off, adj)
    bs, bd = O.uniform_pairs(n, 400, seed=3)
    # add copies of existing edges and in-batch repeats
    bs = np.concatenate([bs, s0[:50], bs[:20]]).astype(np.int32)
    bd = np.concatenate([bd, d0[:50], bd[:20]]).astype(np.int32)
    o2, a2, added = O.insert(off, adj, bs, bd, dedup_batch=True, dedup_graph=True)
    exp = set(zip(s0.tolist(), d0.tolist())) | set(zip(bs.tolist(), bd.tolist()))
    got = multiset(o2, a2)
    assert set(got) == exp and max(got.values()) == 1
    assert added == len(exp) - adj.size
    # dedup_graph only: the FIRST batch copy of an existing edge is dropped
    o3, a3, added3 = O.insert(off, adj, np.array([s0[0]] * 3, np.int32), np.array([d0[0]] * 3, np.int32),
                              dedup_graph=True)
    assert added3 == 2 and multiset(o3, a3)[(int(s0[0]), int(d0[0]))] == 3


def test_erase_multiset_semantics():
    # vertex 0 stores 1 twice (multigraph after an insert without dedup)
    off = np.array([0, 3, 4, 4], np.int64)
    adj = np.array([1, 1, 2, 0], np.int32)
    o2, a2, rem = O.erase(off, adj, np.array([0, 0, 0, 1], np.int32), np.array([1, 1, 1, 2], np.int32))
    # three copies requested, two stored -> both removed; (1,2) absent -> ignored
    assert rem == 2 and o2.tolist() == [0, 1, 2, 2] and a2.tolist() == [2, 0]
    o3, a3, rem3 = O.erase(off, adj, np.array([0, 0], np.int32), np.array([1, 1], np.int32),
                           dedup_batch=True)
    assert rem3 == 1 and a3.tolist() == [1, 2, 0]


def test_insert_self_loops_and_empty_batch(rmat10):
    off, adj = rmat10
    o2, a2, added = O.insert(off, adj, np.array([5], np.int32), np.array([5], np.int32))
    assert added == 1 and 5 in a2[o2[5]:o2[6]]
    o3, a3, added3 = O.insert(off, adj, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert added3 == 0 and np.array_equal(o3, off) and np.array_equal(a3, adj)


# ---------------------------------------------------------------- generators
def test_rmat_generator_invariants():
    s, d = O.rmat_pairs(12, seed=1)
    assert s.size == 16 << 12 and s.min() >= 0 and s.max() < 4096 and d.max() < 4096
    # slicing is consistent (counter-based)
    s2, d2 = O.rmat_pairs(12, seed=1, first=1000, count=77)
    assert np.array_equal(s2, s[1000:1077]) and np.array_equal(d2, d[1000:1077])
    # the top bit follows the quadrant probabilities: P(src top) = c+d = .24, P(dst top) = b+d = .24
    top_s = (s >> 11).mean()
    top_d = (d >> 11).mean()
    assert abs(top_s - 0.24) < 0.01 and abs(top_d - 0.24) < 0.01
    both = ((s >> 11) & (d >> 11)).mean()
    assert abs(both - 0.05) < 0.005
    # different seeds differ
    s3, _ = O.rmat_pairs(12, seed=2)
    assert not np.array_equal(s, s3)


def test_rmat_graph_shape():
    off, adj = O.rmat_graph(16, seed=1)
    assert off.size == 65537
    assert int(off[-1]) == 1819292         # fixed by the generator spec
    s, d = coo(off, adj)
    assert not np.any(s == d)
    key = (s.astype(np.int64) << 32) | d
    assert np.all(np.diff(key) > 0)        # sorted, no duplicates
    rk = (d.astype(np.int64) << 32) | s
    assert np.array_equal(np.sort(rk), key)  # symmetric


def test_uniform_pairs_repeat_rate():
    s, d = O.uniform_pairs(1 << 20, 200000, seed=4)
    keys = (s.astype(np.int64) << 32) | d
    dup = keys.size - np.unique(keys).size
    assert 0.007 < dup / keys.size < 0.013


def test_chung_lu_generator_invariants():
    n = 1 << 14
    cum = O.chung_lu_table(n, 2.1)
    w = np.diff(cum).astype(np.float64)
    # weights follow (i+1)^(-1/1.1) up to the floor
    i = np.array([0, 9, 99, 999])
    assert np.allclose(w[i], np.floor(2**24 * (i + 1.0) ** (-1 / 1.1)))
    s, d = O.chung_lu_pairs(n, 16 * n, seed=9)
    assert s.min() >= 0 and s.max() < n
    # endpoint frequency tracks the weights: vertex 0 vs the expected share
    share = w[0] / w.sum()
    f0 = ((s == 0).sum() + (d == 0).sum()) / (2 * s.size)
    assert abs(f0 - share) < 0.1 * share
    # slicing is consistent
    s2, _ = O.chung_lu_pairs(n, 16 * n, seed=9)
    assert np.array_equal(s, s2)


def test_pick_sources():
    off, adj = O.rmat_graph(10, seed=7)
    src = O.pick_sources(off, 16, seed=3)
    deg = np.diff(off)
    assert src.size == 16 and np.all(deg[src] > 0)


def test_two_hop_vs_networkx(rmat10):
    off, adj = rmat10
    G = to_nx(off, adj)
    for v in [0, 5, O.max_degree_id(off)]:
        assert O.two_hop(off, adj, v) == len(nx.ego_graph(G, v, radius=2))


# ------------------------------------------------------------ edge meta-data
def _list_model_insert(rows, bs, bd, bw, dedup_batch, dedup_graph):
    """rows: dict u -> list of (v, w) in row order; pure-Python restatement."""
    rows = {u: list(r) for u, r in rows.items()}
    seen = set()
    stored = {(u, v) for u, r in rows.items() for v, _ in r}
    batch = sorted(range(len(bs)), key=lambda i: (bs[i], bd[i], i))
    added = 0
    for i in batch:
        k = (int(bs[i]), int(bd[i]))
        first = k not in seen
        seen.add(k)
        if dedup_batch and not first:
            continue
        if dedup_graph and first and k in stored:   # the first batch copy of a stored pair
            continue
        r = rows.setdefault(k[0], [])
        pos = len(r)
        while pos > 0 and r[pos - 1][0] > k[1]:
            pos -= 1
        r.insert(pos, (k[1], int(bw[i])))
        added += 1
    return rows, added


def _rows(off, adj, w):
    return {u: [(int(adj[j]), int(w[j])) for j in range(off[u], off[u + 1])] for u in range(off.size - 1)}


@pytest.mark.parametrize("seed", range(6))
def test_meta_oracle_matches_list_model(seed):
    rng = np.random.default_rng(seed)
    n = 12
    s = rng.integers(0, n, 60)
    d = rng.integers(0, n, 60)
    w = rng.integers(0, 1000, 60)
    o0, a0, p0 = _coo_csr(n, s, d)
    off, adj, (mw,) = O.csr_meta(o0, a0, [w[p0]])
    rows = {u: [] for u in range(n)}
    for i in sorted(range(60), key=lambda i: (s[i], d[i], i)):
        rows[int(s[i])].append((int(d[i]), int(w[i])))
    assert _rows(off, adj, mw) == rows
    bs, bd, bw = rng.integers(0, n, 40), rng.integers(0, n, 40), rng.integers(0, 1000, 40)
    for db in (False, True):
        for dg in (False, True):
            o1, a1, (w1,), added = O.insert_meta(off, adj, [mw], bs, bd, [bw], db, dg)
            exp, eadded = _list_model_insert(rows, bs, bd, bw, db, dg)
            assert added == eadded and _rows(o1, a1, w1) == exp
            # the edge multiset agrees with the pinned insert model (fields dropped)
            oo, aa, _ = O.insert(off, adj, bs.astype(np.int32), bd.astype(np.int32), db, dg)
            assert np.array_equal(oo, o1) and np.array_equal(aa, a1)
    es, ed = bs[:25], bd[:25]
    o2, a2, (w2,), removed = O.erase_meta(off, adj, [mw], es, ed)
    oo, aa, rem = O.erase(off, adj, es.astype(np.int32), ed.astype(np.int32))
    assert removed == rem and np.array_equal(oo, o2) and np.array_equal(aa, a2)
    # oldest copies go first: the survivors of each (u, v) are its newest copies
    for u in range(n):
        for v in set(adj[off[u]:off[u + 1]].tolist()):
            before = [x for (y, x) in rows[u] if y == v]
            after = [int(w2[j]) for j in range(o2[u], o2[u + 1]) if a2[j] == v]
            assert after == before[len(before) - len(after):]


def _coo_csr(n, s, d):
    p = np.lexsort((np.arange(s.size), d, s))
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off), d[p].astype(np.int32), p


def test_meta_batch_duplicates_keep_first():
    off = np.array([0, 0, 0, 0], np.int64)
    adj = np.array([], np.int32)
    o, a, (w,), added = O.insert_meta(off, adj, [np.array([], np.int64)], [1, 1, 0, 1], [2, 2, 1, 2],
                                      [np.array([7, 8, 9, 10])], dedup_batch=True)
    assert added == 2 and a.tolist() == [1, 2] and w.tolist() == [9, 7]
    o, a, (w,), added = O.insert_meta(off, adj, [np.array([], np.int64)], [1, 1], [2, 2],
                                      [np.array([7, 8])])
    assert a.tolist() == [2, 2] and w.tolist() == [7, 8]


# ------------------------------------------------- hand-derived rule pins
def test_tc2_credit_rule_hand_case():
    """TriangleCounting2 credit rule on a stored directed graph, worked out by
    hand from primitives/Operator++.i.cuh:120-131 and triangle2.cu:121: skip
    src == dst and deg(src) < 2; credit |N(s) ∩ N(d)| to the endpoint of
    smaller degree, ties to the destination.

    rows: 0:[1,2] 1:[2,3] 2:[3] 3:[3] 4:[4,5] 5:[]
      (0,1) deg 2/2 tie -> 1 gets |{1,2}∩{2,3}| = 1
      (0,2) 2 > 1 -> 2 gets 0;  (1,2) -> 2 gets 1;  (1,3) -> 3 gets 1
      (2,3) deg(2) = 1 -> skipped (else 3 would get 1 more)
      (3,3), (4,4) src == dst -> skipped (else 4 would get 2)
      (4,5) -> 5 gets 0
    Ties to the source would give [1,0,1,1,0,0]; crediting the larger
    endpoint [*,2,0,0,..]."""
    off = np.array([0, 2, 4, 5, 6, 8, 8], np.int64)
    adj = np.array([1, 2, 2, 3, 3, 3, 4, 5], np.int32)
    assert O.tc2_per_vertex(off, adj).tolist() == [0, 1, 1, 1, 0, 0]


def test_directed_by_degree_equal_degree_tie():
    """GraphStd.cpp:265 keeps u->v iff deg u < deg v, or deg u == deg v and
    u < v.  Cycle 0-1-2-3-0 plus the pendant 4-2 (degrees 2,2,3,2,1), by hand:
    0->1 and 0->3 (ties, lower id first), 1->2, 3->2, 4->2."""
    pairs = [(0, 1), (1, 2), (2, 3), (3, 0), (2, 4)]
    s = np.array([p[0] for p in pairs], np.int32)
    d = np.array([p[1] for p in pairs], np.int32)
    off, adj = O.coo_to_csr(5, s, d)
    assert np.diff(off).tolist() == [2, 2, 3, 2, 1]
    do, da = O.directed_by_degree(off, adj)
    assert do.tolist() == [0, 2, 3, 3, 4, 5] and da.tolist() == [1, 3, 2, 2, 2]


def test_gmtx_max_degree_tie_hand():
    """example/G.mtx read UNDIRECTED (0-based): degrees worked out by hand from
    its 12 arcs are [3,1,4,4,3,1,2,2,2]; vertices 2 and 3 tie at 4 and
    std::max_element takes the first (GraphStd.i.hpp:296-299): 2."""
    n, s, d, _ = O.read_market(os.path.join(GOLD, "G.mtx"))
    off, adj = O.coo_to_csr(n, s, d, symmetrize=True, drop_self_loops=False, dedup=True)
    assert np.diff(off).tolist() == [3, 1, 4, 4, 3, 1, 2, 2, 2]
    assert O.max_degree_id(off) == 2


# --------------------------------- parallel checker == sequential checker
@pytest.mark.parametrize("threads", [2, 5])
def test_parallel_checker_matches_sequential(threads):
    s1, d1 = O.rmat_pairs(13, seed=5)
    s2, d2 = O.rmat_pairs(13, seed=5, threads=threads)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)
    n = 1 << 13
    # keep some self loops and repeats in the input
    s1 = np.concatenate([s1, np.arange(50, dtype=np.int32), s1[:1000]])
    d1 = np.concatenate([d1, np.arange(50, dtype=np.int32), d1[:1000]])
    for sym in (True, False):
        for drop in (True, False):
            a = O.coo_to_csr(n, s1, d1, symmetrize=sym, drop_self_loops=drop)
            b = O.coo_to_csr(n, s1, d1, symmetrize=sym, drop_self_loops=drop, threads=threads)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    off, adj = O.rmat_graph(13, seed=5)
    o2, a2 = O.rmat_graph(13, seed=5, threads=threads)
    assert np.array_equal(off, o2) and np.array_equal(adj, a2)
    do, da = O.directed_by_degree(off, adj)
    assert O.host_tc(do, da) == O.host_tc(do, da, threads=threads)
    assert O.triangles(off, adj) == O.triangles(off, adj, threads=threads)
    cs, cd = O.chung_lu_pairs(1 << 12, 1 << 15, seed=3)
    cs2, cd2 = O.chung_lu_pairs(1 << 12, 1 << 15, seed=3, threads=threads)
    assert np.array_equal(cs, cs2) and np.array_equal(cd, cd2)
    srcs = O.pick_sources(off, 6, seed=1)
    many = O.bfs_many(off, adj, srcs, threads=threads)
    for s, (dist, sc) in zip(srcs, many):
        ed, esc = O.bfs(off, adj, int(s))
        assert np.array_equal(dist, ed) and sc == esc
