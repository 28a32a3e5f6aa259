"""GPU parity: triangle counting (TriangleCounting2) against the oracle."""
import json
import os
from math import comb

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


def tc_count(hn, h):
    tc = hn.TriangleCounting2(h)
    tc.init()
    tc.run_undirected()
    return tc.countTriangles(), tc.copyTCToHost()


def tc_stored(hn, h):
    tc = hn.TriangleCounting2(h)
    tc.init()
    tc.run()
    return tc.countTriangles(), tc.copyTCToHost()


@pytest.mark.parametrize("scale,seed", [(10, 7), (14, 2), (16, 1)])
def test_tc_rmat(hn, scale, seed):
    print(f"case tc rmat scale={scale} seed={seed}", flush=True)
    off, adj = O.rmat_graph(scale, seed)
    h = hn.Hornet.rmat(scale, seed)
    total, per = tc_count(hn, h)
    assert total == O.triangles(off, adj)
    do, da = O.directed_by_degree(off, adj)
    assert np.array_equal(per, O.tc2_per_vertex(do, da))


def test_tc2_on_oriented_storage(hn):
    # the reference driver's form: the Hornet holds the DIRECTED_BY_DEGREE graph
    print("case tc2 oriented storage", flush=True)
    off, adj = O.rmat_graph(14, 3)
    do, da = O.directed_by_degree(off, adj)
    h = hn.Hornet.from_csr(do, da)
    total, per = tc_stored(hn, h)
    assert total == O.triangles(off, adj)
    assert np.array_equal(per, O.tc2_per_vertex(do, da))


def test_tc2_on_symmetric_storage(hn):
    # every triangle is seen from its six arcs
    print("case tc2 symmetric storage", flush=True)
    off, adj = O.rmat_graph(12, 4)
    h = hn.Hornet.from_csr(off, adj)
    total, per = tc_stored(hn, h)
    assert total == 6 * O.triangles(off, adj)
    assert np.array_equal(per, O.tc2_per_vertex(off, adj))


@pytest.mark.parametrize("name", ["G.mtx", "G_cc2.mtx", "rome99.gr"])
def test_tc_fixtures(hn, name):
    print(f"case tc fixture {name}", flush=True)
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)[name]
    path = os.path.join(GOLD, name)
    if name.endswith(".mtx"):
        n, s, d, _ = O.read_market(path)
    else:
        n, s, d = O.read_dimacs9(path)
    off, adj = O.coo_to_csr(n, s, d, symmetrize=True, drop_self_loops=True, dedup=True)
    h = hn.Hornet.from_csr(off, adj)
    total, _ = tc_count(hn, h)
    assert total == g["triangles"]


def test_tc_complete_graph_heavy_class(hn):
    # K_600: oriented out-degrees reach 599, so intersections use the
    # 16-lane and whole-wave classes
    print("case tc K600", flush=True)
    n = 600
    s, d = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    mask = s != d
    off, adj = O.coo_to_csr(n, s[mask].astype(np.int32), d[mask].astype(np.int32))
    h = hn.Hornet.from_csr(off, adj)
    total, per = tc_count(hn, h)
    assert total == comb(n, 3)
    do, da = O.directed_by_degree(off, adj)
    assert np.array_equal(per, O.tc2_per_vertex(do, da))
    total2, per2 = tc_stored(hn, h)
    assert total2 == 6 * comb(n, 3)
    assert np.array_equal(per2, O.tc2_per_vertex(off, adj))


def test_tc_no_triangles_and_empty(hn):
    print("case tc trivial", flush=True)
    # a path has no triangles
    off = np.array([0, 1, 3, 5, 6], np.int64)
    adj = np.array([1, 0, 2, 1, 3, 2], np.int32)
    total, per = tc_count(hn, hn.Hornet.from_csr(off, adj))
    assert total == 0 and not per.any()
    total, _ = tc_count(hn, hn.Hornet.from_csr(np.zeros(5, np.int64), np.zeros(0, np.int32)))
    assert total == 0


def test_tc2_credit_rule_hand_case(hn):
    """The hand-worked stored graph of the oracle pin (primitives/Operator++.i.cuh:120-131,
    triangle2.cu:121): smaller-degree endpoint credited, ties to the destination,
    src == dst and deg(src) < 2 skipped -> per-vertex [0, 1, 1, 1, 0, 0]."""
    print("case tc2 hand-worked credit rule", flush=True)
    off = np.array([0, 2, 4, 5, 6, 8, 8], np.int64)
    adj = np.array([1, 2, 2, 3, 3, 3, 4, 5], np.int32)
    h = hn.Hornet.from_csr(off, adj)
    total, per = tc_stored(hn, h)
    assert per.tolist() == [0, 1, 1, 1, 0, 0] and total == 3


def test_gmtx_max_degree_tie_hand(hn):
    """G.mtx read UNDIRECTED: vertices 2 and 3 tie at degree 4; the lowest id, 2,
    is Hornet::max_degree_id (hand value, GraphStd.i.hpp:296-299)."""
    print("case G.mtx max-degree tie", flush=True)
    n, s, d, _ = O.read_market(os.path.join(GOLD, "G.mtx"))
    off, adj = O.coo_to_csr(n, s, d, symmetrize=True, drop_self_loops=False, dedup=True)
    h = hn.Hornet.from_csr(off, adj)
    assert h.max_degree_id() == 2 and h.max_degree() == 4
