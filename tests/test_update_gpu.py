"""GPU parity: BatchUpdate insert / erase against the host set model."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


def coo(off, adj):
    s = np.repeat(np.arange(off.size - 1), np.diff(off)).astype(np.int32)
    return s, adj.astype(np.int32)


def check_equal(h, off, adj):
    s, d = h.getCOO()
    xs, xd = coo(off, adj)
    assert h.nE() == adj.size
    assert np.array_equal(s.cpu().numpy(), xs)
    assert np.array_equal(d.cpu().numpy(), xd)
    assert np.array_equal(h.degrees().cpu().numpy(), np.diff(off))


@pytest.fixture()
def g12(hn):
    off, adj = O.rmat_graph(12, seed=5)
    return off, adj, hn.Hornet.from_csr(off, adj)


@pytest.mark.parametrize("dedup_batch,dedup_graph", [(False, False), (True, False), (False, True), (True, True)])
def test_insert_flags(hn, g12, dedup_batch, dedup_graph):
    print(f"case insert dedup_batch={dedup_batch} dedup_graph={dedup_graph}", flush=True)
    off, adj, h = g12
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 3000, seed=8)
    s0, d0 = coo(off, adj)
    # existing edges (some twice) and self loops
    bs = np.concatenate([bs, s0[::97], s0[::194], np.arange(10)]).astype(np.int32)
    bd = np.concatenate([bd, d0[::97], d0[::194], np.arange(10)]).astype(np.int32)
    got = h.insert(hn.BatchUpdate(bs, bd), dedup_batch, dedup_graph)
    o2, a2, added = O.insert(off, adj, bs, bd, dedup_batch, dedup_graph)
    assert got == added
    check_equal(h, o2, a2)


@pytest.mark.parametrize("dedup", [False, True])
def test_erase(hn, g12, dedup):
    print(f"case erase dedup={dedup}", flush=True)
    off, adj, h = g12
    s0, d0 = coo(off, adj)
    rng = np.random.default_rng(1)
    pick = rng.choice(adj.size, 2500, replace=False)
    bs, bd = s0[pick], d0[pick]
    # repeats, absent edges, and a whole adjacency list
    v = int(np.argmax(np.diff(off)))
    bs = np.concatenate([bs, bs[:300], np.array([1, 2, 3], np.int32), np.full(off[v + 1] - off[v], v, np.int32)])
    bd = np.concatenate([bd, bd[:300], np.array([1, 2, 3], np.int32), adj[off[v]:off[v + 1]]])
    got = h.erase(hn.BatchUpdate(bs.astype(np.int32), bd.astype(np.int32)), dedup)
    o2, a2, removed = O.erase(off, adj, bs, bd, dedup)
    assert got == removed
    check_equal(h, o2, a2)
    assert h.degrees().cpu().numpy()[v] == 0


def test_multigraph_erase_multiplicity(hn):
    print("case multigraph erase", flush=True)
    off = np.array([0, 2, 3, 3], np.int64)
    adj = np.array([1, 2, 0], np.int32)
    h = hn.Hornet.from_csr(off, adj)
    h.insert(hn.BatchUpdate(np.array([0, 0, 0], np.int32), np.array([1, 1, 2], np.int32)))
    o1, a1, _ = O.insert(off, adj, np.array([0, 0, 0], np.int32), np.array([1, 1, 2], np.int32))
    check_equal(h, o1, a1)   # 0 -> [1,1,1,2,2]
    bs, bd = np.array([0, 0, 0, 0], np.int32), np.array([1, 1, 2, 2], np.int32)
    assert h.erase(hn.BatchUpdate(bs, bd)) == 4
    o2, a2, _ = O.erase(o1, a1, bs, bd)
    check_equal(h, o2, a2)   # 0 -> [1]


def test_rounds_with_bfs_and_tc(hn):
    # C3 in miniature: rounds of symmetric inserts (with repeats) and deletes
    # sampled from the current undirected edge set, edge set + TC + BFS each round
    off, adj = O.rmat_graph(14, seed=24)
    h = hn.Hornet.from_csr(off, adj)
    n = off.size - 1
    for r in range(4):
        print(f"case rounds r={r}", flush=True)
        us, ud = O.uniform_pairs(n, 4000, seed=4 + r)
        keep = us != ud
        bs = np.concatenate([us[keep], ud[keep]])
        bd = np.concatenate([ud[keep], us[keep]])
        h.insert(hn.BatchUpdate(bs, bd), True, True)
        off, adj, _ = O.insert(off, adj, bs, bd, True, True)
        check_equal(h, off, adj)
        s0, d0 = coo(off, adj)
        und = np.flatnonzero(s0 < d0)
        pick = und[np.random.default_rng(r).integers(0, und.size, 4000)]
        es = np.concatenate([s0[pick], d0[pick]])
        ed = np.concatenate([d0[pick], s0[pick]])
        h.erase(hn.BatchUpdate(es, ed), True)
        off, adj, _ = O.erase(off, adj, es, ed, True)
        check_equal(h, off, adj)
        tc = hn.TriangleCounting2(h)
        tc.init()
        tc.run_undirected()
        assert tc.countTriangles() == O.triangles(off, adj)
        src = O.max_degree_id(off)
        bfs = hn.BfsTopDown(h)
        bfs.set_parameters(src)
        bfs.run()
        assert np.array_equal(bfs.distances().cpu().numpy(), O.bfs(off, adj, src)[0])


def test_invalid_batch_leaves_graph_unchanged(hn, g12):
    print("case invalid batch", flush=True)
    from hornet_amd import HornetError
    off, adj, h = g12
    with pytest.raises(HornetError):
        h.insert(hn.BatchUpdate(np.array([0, 5], np.int32), np.array([1, off.size + 3], np.int32)))
    with pytest.raises(HornetError):
        h.erase(hn.BatchUpdate(np.array([-1], np.int32), np.array([1], np.int32)))
    check_equal(h, off, adj)


def test_arena_exhaustion_is_reported_and_harmless(hn):
    print("case arena exhaustion", flush=True)
    from hornet_amd import HornetError
    off, adj = O.rmat_graph(12, seed=9)
    h = hn.Hornet.from_csr(off, adj, slack=1.0)
    used0, cap = h.memory()
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 400000, seed=1)
    with pytest.raises(HornetError) as ei:
        h.insert(hn.BatchUpdate(bs, bd))
    assert ei.value.status == 2
    check_equal(h, off, adj)
    assert h.memory() == (used0, cap)
    # a batch that fits still works afterwards
    bs2, bd2 = O.uniform_pairs(n, 50, seed=2)
    h.insert(hn.BatchUpdate(bs2, bd2))
    o2, a2, _ = O.insert(off, adj, bs2, bd2)
    check_equal(h, o2, a2)


def test_free_lists_recycle_blocks(hn):
    print("case free-list reuse", flush=True)
    off, adj = O.rmat_graph(12, seed=3)
    h = hn.Hornet.from_csr(off, adj, slack=1.5)
    s0, d0 = coo(off, adj)
    for r in range(20):
        h.erase(hn.BatchUpdate(s0[r::20], d0[r::20]))
        h.insert(hn.BatchUpdate(s0[r::20], d0[r::20]))
    check_equal(h, off, adj)
    used, cap = h.memory()
    assert used <= cap


def test_host_pointer_batches(hn, g12):
    print("case host pointers", flush=True)
    from hornet_amd import _lib
    off, adj, h = g12
    bs, bd = O.uniform_pairs(off.size - 1, 100, seed=77)
    out = ctypes.c_int64()
    _lib.call("hn_graph_insert", h.handle, bs.ctypes.data_as(ctypes.c_void_p), bd.ctypes.data_as(ctypes.c_void_p),
              bs.size, 0, ctypes.byref(out), None)
    o2, a2, added = O.insert(off, adj, bs, bd)
    assert out.value == added
    check_equal(h, o2, a2)
    _lib.call("hn_graph_erase", h.handle, bs.ctypes.data_as(ctypes.c_void_p), bd.ctypes.data_as(ctypes.c_void_p),
              bs.size, 0, ctypes.byref(out), None)
    check_equal(h, off, adj)


def test_insert_erase_roundtrip_large(hn):
    # size-independent property at a larger scale: insert without dedup is a
    # multiset union and erase without dedup a multiset difference, so
    # inserting a batch and erasing the same batch restores the graph exactly
    print("case roundtrip s20", flush=True)
    import torch
    h = hn.Hornet.rmat(20, seed=2)
    s0, d0 = h.getCOO()
    n = h.nV()
    us, ud = hn.uniform_pairs_device(n, 1 << 20, seed=11)
    bs = torch.cat([us, ud])
    bd = torch.cat([ud, us])
    added = h.insert(hn.BatchUpdate(bs, bd))
    assert added == bs.numel() and h.nE() == s0.numel() + added
    removed = h.erase(hn.BatchUpdate(bs, bd))
    assert added == removed
    s1, d1 = h.getCOO()
    assert torch.equal(s0, s1) and torch.equal(d0, d1)


def test_empty_batches(hn, g12):
    off, adj, h = g12
    assert h.insert(hn.BatchUpdate(np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
    assert h.erase(hn.BatchUpdate(np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
    check_equal(h, off, adj)
