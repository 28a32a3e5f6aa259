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


def test_arena_grows_on_demand(hn):
    """An insert that does not fit the arena grows it and runs again
    (BlockArrayManager adds BlockArrays, BlockArray.i.cuh:138-161)."""
    print("case arena growth", flush=True)
    off, adj = O.rmat_graph(12, seed=9)
    h = hn.Hornet.from_csr(off, adj, slack=1.0)
    _, cap0 = h.memory()
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 400000, seed=1)
    added = h.insert(hn.BatchUpdate(bs, bd))
    o2, a2, e_added = O.insert(off, adj, bs, bd)
    assert added == e_added
    check_equal(h, o2, a2)
    used, cap = h.memory()
    assert cap > cap0 and used <= cap
    # a batch that fits still works afterwards
    bs2, bd2 = O.uniform_pairs(n, 50, seed=2)
    h.insert(hn.BatchUpdate(bs2, bd2))
    o3, a3, _ = O.insert(o2, a2, bs2, bd2)
    check_equal(h, o3, a3)


def test_empty_graph_grows_past_ten_times_its_arena(hn):
    """Hornet(nV) with no edges (4096 units = 131072 slots) filled with an
    R-MAT s15 graph and more pairs, in batches, through several growths; the
    batched BFS (whose captured level step reads the arena) stays exact."""
    print("case empty graph growth", flush=True)
    n = 1 << 15
    h = hn.Hornet.from_csr(np.zeros(n + 1, np.int64), np.zeros(0, np.int32))
    _, cap0 = h.memory()
    off = np.zeros(n + 1, np.int64)
    adj = np.zeros(0, np.int32)
    gs, gd = O.rmat_pairs(15, seed=4)
    us, ud = O.uniform_pairs(n, 500000, seed=5)
    ms = hn.BfsMulti(h)
    srcs = None
    for i, (bs, bd) in enumerate([(gs[:100000], gd[:100000]), (gs[100000:], gd[100000:]), (us, ud),
                                  (gd, gs)]):
        h.insert(hn.BatchUpdate(bs, bd), True, True)
        off, adj, _ = O.insert(off, adj, bs, bd, True, True)
        check_equal(h, off, adj)
        if srcs is None:
            srcs = O.pick_sources(off, 8, seed=1)
        dist = ms.run(srcs).cpu().numpy()
        for j, s in enumerate(srcs):
            assert np.array_equal(dist[j], O.bfs(off, adj, int(s))[0]), f"batch {i} source {s}"
    used, cap = h.memory()
    assert adj.size >= 10 * cap0 and cap >= used >= adj.size


def test_failed_batch_returns_its_blocks(hn, monkeypatch):
    """A batch that fails after its blocks were allocated (a device allocation
    failure, forced from the host) hands them back: the arena bookkeeping and
    the graph are as before, and the next batch works."""
    print("case failed batch rollback", flush=True)
    from hornet_amd import HornetError
    off, adj = O.rmat_graph(12, seed=9)
    h = hn.Hornet.from_csr(off, adj, slack=1.5)
    mem0 = h.memory()
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 20000, seed=3)
    s0, d0 = coo(off, adj)
    es, ed = s0[::2], d0[::2]
    monkeypatch.setenv("HORNET_HIP_FAIL_SIDE_RESERVE", "1")
    with pytest.raises(HornetError) as ei:
        h.insert(hn.BatchUpdate(bs, bd))
    assert ei.value.status == 2
    with pytest.raises(HornetError) as ei:
        h.erase(hn.BatchUpdate(es, ed))
    assert ei.value.status == 2
    monkeypatch.delenv("HORNET_HIP_FAIL_SIDE_RESERVE")
    assert h.memory() == mem0
    check_equal(h, off, adj)
    h.insert(hn.BatchUpdate(bs, bd))
    o2, a2, _ = O.insert(off, adj, bs, bd)
    check_equal(h, o2, a2)
    h.erase(hn.BatchUpdate(es, ed))
    o3, a3, _ = O.erase(o2, a2, es, ed)
    check_equal(h, o3, a3)


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


def presence_symmetric(off, adj):
    """(u,w) stored (any number of copies) iff (w,u) stored."""
    s, d = coo(off, adj)
    n = off.size - 1
    k = np.unique(s.astype(np.int64) * n + d)
    rk = np.unique(d.astype(np.int64) * n + s)
    return k.size == rk.size and np.array_equal(k, rk)


def multigraph_base():
    """R-MAT 2^10, symmetric, with 2 copies of every pair (u,w), u < w, u % 5 == 0."""
    n = 1 << 10
    s, d = O.rmat_pairs(10, seed=12)
    keep = s != d
    s, d = s[keep], d[keep]
    lo, hi = np.minimum(s, d), np.maximum(s, d)
    k = np.unique(lo.astype(np.int64) * n + hi)
    lo, hi = (k // n).astype(np.int32), (k % n).astype(np.int32)
    twice = lo % 5 == 0
    s = np.concatenate([lo, lo[twice]])
    d = np.concatenate([hi, hi[twice]])
    return O.coo_to_csr(n, s, d, symmetrize=True, dedup=False)


def sym_cases(off, adj):
    """(name, op, src, dst, flags, symmetric after) over a symmetric multigraph."""
    n = off.size - 1
    s, d = coo(off, adj)
    fwd = s < d
    key = s.astype(np.int64) * n + d
    u, c = np.unique(key[fwd], return_counts=True)
    one = u[c == 1]
    two = u[c == 2]
    a1, b1 = (one // n).astype(np.int32), (one % n).astype(np.int32)
    a2, b2 = (two // n).astype(np.int32), (two % n).astype(np.int32)
    assert a1.size > 200 and a2.size > 50
    stored = set(key.tolist())
    rng = np.random.default_rng(9)
    x = rng.integers(0, n, 400).astype(np.int32)
    y = rng.integers(0, n, 400).astype(np.int32)
    new = (x != y) & np.array([int(p) * n + int(q) not in stored and int(q) * n + int(p) not in stored
                              for p, q in zip(x, y)])
    x, y = x[new], y[new]
    cat = np.concatenate
    return [
        ("erase both directions", "erase", cat([a1[:100], b1[:100]]), cat([b1[:100], a1[:100]]), (False,), True),
        ("erase one direction", "erase", a1[:50], b1[:50], (False,), False),
        ("erase absent pairs, one direction", "erase", cat([x[:50], a1[:20], b1[:20]]),
         cat([y[:50], b1[:20], a1[:20]]), (False,), True),
        ("erase one of two copies, both directions", "erase", cat([a2[:30], b2[:30]]), cat([b2[:30], a2[:30]]),
         (False,), True),
        ("erase one of two copies, one direction", "erase", a2[:30], b2[:30], (False,), True),
        ("erase both copies one way, one the other way", "erase", cat([a2[:5], a2[:5], b2[:5]]),
         cat([b2[:5], b2[:5], a2[:5]]), (False,), False),
        ("erase both copies one way, deduplicated batch", "erase", cat([a2[:5], a2[:5], b2[:5]]),
         cat([b2[:5], b2[:5], a2[:5]]), (True,), True),
        ("insert new pairs, both directions", "insert", cat([x[:60], y[:60]]), cat([y[:60], x[:60]]),
         (True, True), True),
        ("insert new pairs, one direction", "insert", x[:60], y[:60], (True, True), False),
        ("insert stored pairs, one direction", "insert", a1[:60], b1[:60], (False, False), True),
        ("insert stored pairs, graph-deduplicated", "insert", a1[:60], b1[:60], (False, True), True),
    ]


def test_symmetry_tracked_by_each_batch(hn):
    """After each batch on a graph known to be symmetric, the symmetry flag
    (seen through BfsMulti, which pulls only over a symmetric adjacency) equals
    the symmetry of the oracle's adjacency."""
    off0, adj0 = multigraph_base()
    assert presence_symmetric(off0, adj0)
    srcs = O.pick_sources(off0, 8, seed=2)
    for name, op, bs, bd, flags, sym in sym_cases(off0, adj0):
        print(f"case symmetry: {name}", flush=True)
        h = hn.Hornet.from_csr(off0, adj0)
        ms = hn.BfsMulti(h)
        ms.set_params(0, 1 << 20)   # pull at every level the adjacency allows
        ms.run(srcs)
        assert ms.stats.pull_levels >= 1
        if op == "erase":
            h.erase(hn.BatchUpdate(bs, bd), *flags)
            off, adj, _ = O.erase(off0, adj0, bs, bd, *flags)
        else:
            h.insert(hn.BatchUpdate(bs, bd), *flags)
            off, adj, _ = O.insert(off0, adj0, bs, bd, *flags)
        assert presence_symmetric(off, adj) == sym
        check_equal(h, off, adj)
        got = ms.run(srcs).cpu().numpy()
        assert np.array_equal(got, np.stack([O.bfs(off, adj, int(v))[0] for v in srcs]))
        assert (ms.stats.pull_levels >= 1) == sym, name


def hub_graph(seed):
    """2^16 vertices: four hubs of ~30000 entries (capacity 32768) over uniform
    background pairs, directed, with repeated pairs (a multigraph)."""
    rng = np.random.default_rng(seed)
    n = 1 << 16
    hubs = [5, 77, 1000, 40000]
    s = [np.full(30000, v, np.int32) for v in hubs]
    d = [rng.choice(n, 30000, replace=False).astype(np.int32) for _ in hubs]
    bs, bd = O.uniform_pairs(n, 60000, seed=seed)
    s = np.concatenate(s + [bs, bs[:5000]])
    d = np.concatenate(d + [bd, bd[:5000]])
    off, adj = O.coo_to_csr(n, s, d, symmetrize=False, drop_self_loops=False, dedup=False)
    return off, adj, hubs


def hub_erase_batch(off, adj, hubs, seed, frac):
    """A fraction of every hub row (40 % of 30000, then 8 % of 18000: the rows
    stay in their 32768-entry blocks and are compacted in place, with shifts of
    thousands of slots) and a seventh of the other entries."""
    rng = np.random.default_rng(seed)
    n = off.size - 1
    s0, _ = coo(off, adj)
    pick = np.zeros(adj.size, bool)
    pick[::7] = True
    for v in hubs:
        deg = int(off[v + 1] - off[v])
        pick[off[v]:off[v + 1]] = False
        pick[off[v] + rng.choice(deg, int(deg * frac), replace=False)] = True
    idx = rng.permutation(np.nonzero(pick)[0])
    return s0[idx].astype(np.int32), adj[idx].astype(np.int32)


def test_erase_hub_rows_in_place(hn):
    """Long rows compacted in place over many copy tiles, where most tiles start
    behind a deleted range of their row; two rounds, the second on the
    compacted rows."""
    print("case erase hub rows in place", flush=True)
    off, adj, hubs = hub_graph(3)
    h = hn.Hornet.from_csr(off, adj)
    for r in range(2):
        es, ed = hub_erase_batch(off, adj, hubs, 10 + r, (0.4, 0.08)[r])
        got = h.erase(hn.BatchUpdate(es, ed))
        off, adj, removed = O.erase(off, adj, es, ed)
        assert got == removed
        check_equal(h, off, adj)


def hub_insert_batch(off, hubs, seed, per_hub=2000):
    """per_hub new entries into every hub row (30000 -> 32000 of a 32768-entry
    block: the rows grow in place, their old entries moving up by up to
    per_hub slots across many copy tiles), repeats included, and uniform
    background pairs."""
    rng = np.random.default_rng(seed)
    n = off.size - 1
    s = [np.full(per_hub, v, np.int32) for v in hubs]
    d = [rng.integers(0, n, per_hub).astype(np.int32) for _ in hubs]
    bs, bd = O.uniform_pairs(n, 20000, seed=seed)
    return np.concatenate(s + [bs]), np.concatenate(d + [bd])


@pytest.mark.parametrize("dedup", [False, True])
def test_insert_hub_rows_in_place(hn, dedup):
    print(f"case insert hub rows in place dedup={dedup}", flush=True)
    off, adj, hubs = hub_graph(5)
    h = hn.Hornet.from_csr(off, adj)
    bs, bd = hub_insert_batch(off, hubs, 30)
    got = h.insert(hn.BatchUpdate(bs, bd), dedup, dedup)
    off, adj, added = O.insert(off, adj, bs, bd, dedup, dedup)
    assert got == added
    assert all(30000 < off[v + 1] - off[v] <= 32768 for v in hubs)
    check_equal(h, off, adj)


def test_updates_on_many_short_rows(hn):
    """Rows of 4-8 entries, each losing or gaining about one entry in place:
    a 1024-item copy tile then spans several hundred runs, more than its LDS
    stage holds (257), so the copies search their runs in global memory."""
    print("case updates on many short rows", flush=True)
    rng = np.random.default_rng(17)
    n = 1 << 14
    deg = rng.integers(4, 9, n)
    s = np.repeat(np.arange(n), deg).astype(np.int32)
    d = rng.integers(0, n, s.size).astype(np.int32)
    off, adj = O.coo_to_csr(n, s, d, symmetrize=False, drop_self_loops=False, dedup=True)
    h = hn.Hornet.from_csr(off, adj)
    s0, d0 = coo(off, adj)
    # one stored entry of (almost) every row, from anywhere in the row
    first = off[:-1][np.diff(off) > 0]
    pick = first + (rng.integers(0, 1 << 30, first.size) % np.diff(off)[np.diff(off) > 0])
    es, ed = s0[pick].astype(np.int32), d0[pick].astype(np.int32)
    assert h.erase(hn.BatchUpdate(es, ed)) == es.size
    off, adj, _ = O.erase(off, adj, es, ed)
    check_equal(h, off, adj)
    # one or two new entries per row, interleaved with the stored ones
    bs = np.concatenate([np.arange(n), np.arange(0, n, 2)]).astype(np.int32)
    bd = rng.integers(0, n, bs.size).astype(np.int32)
    got = h.insert(hn.BatchUpdate(bs, bd), True, True)
    off, adj, added = O.insert(off, adj, bs, bd, True, True)
    assert got == added
    check_equal(h, off, adj)


def test_update_stats_match_host_model(hn):
    """hn_graph_update_stats after an insert and an erase: the touched vertices,
    their degree sum and the applied count follow from the batch and the host
    model; relocations are those whose block no longer fits."""
    print("case update stats", flush=True)
    off, adj = O.rmat_graph(12, seed=4)
    h = hn.Hornet.from_csr(off, adj)
    n = off.size - 1
    deg = np.diff(off)
    bs, bd = O.uniform_pairs(n, 3000, seed=8)
    added = h.insert(hn.BatchUpdate(bs, bd), True, True)
    st = h.update_stats()
    o2, a2, e_added = O.insert(off, adj, bs, bd, True, True)
    assert st["kind"] == 1 and st["submitted"] == bs.size and st["applied"] == added == e_added
    # touched = distinct sources of the kept entries
    keep_src = np.repeat(np.arange(n), np.diff(o2) - deg)
    T = np.unique(keep_src)
    assert st["touched"] == T.size and st["touched_deg"] == int(deg[T].sum())
    cap = np.where(deg > 0, np.maximum(32, 2 ** np.ceil(np.log2(np.maximum(deg, 1)))), 0)
    R = T[np.diff(o2)[T] > cap[T]]
    assert st["relocated"] == R.size and st["relocated_deg"] == int(deg[R].sum())
    s2, d2 = coo(o2, a2)
    es, ed = s2[::3], d2[::3]
    removed = h.erase(hn.BatchUpdate(es, ed))
    st = h.update_stats()
    assert st["kind"] == 2 and st["applied"] == removed == es.size
    assert st["touched"] == np.unique(es).size
