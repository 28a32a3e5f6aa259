"""GPU parity: storage, generators and BFS TopDown against the oracle."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
INF = 2**31 - 1


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


def coo(off, adj):
    s = np.repeat(np.arange(off.size - 1), np.diff(off)).astype(np.int32)
    return s, adj.astype(np.int32)


def export(h):
    s, d = h.getCOO()
    return s.cpu().numpy(), d.cpu().numpy()


def load_fixture(name):
    path = os.path.join(GOLD, name)
    if name.endswith(".mtx"):
        n, s, d, _ = O.read_market(path)
        return O.coo_to_csr(n, s, d, symmetrize=True, drop_self_loops=False, dedup=True)
    n, s, d = O.read_dimacs9(path)
    return O.coo_to_csr(n, s, d, symmetrize=False, drop_self_loops=False, dedup=False, sort=True)


@pytest.mark.parametrize("scale,seed", [(8, 1), (12, 5), (16, 1)])
def test_rmat_pairs_match_oracle(hn, scale, seed):
    print(f"case rmat_pairs scale={scale} seed={seed}", flush=True)
    s, d = hn.rmat_pairs_device(scale, seed)
    es, ed = O.rmat_pairs(scale, seed)
    assert np.array_equal(s.cpu().numpy(), es) and np.array_equal(d.cpu().numpy(), ed)


def test_uniform_pairs_match_oracle(hn):
    print("case uniform_pairs", flush=True)
    s, d = hn.uniform_pairs_device(1 << 14, 50000, seed=9)
    es, ed = O.uniform_pairs(1 << 14, 50000, seed=9)
    assert np.array_equal(s.cpu().numpy(), es) and np.array_equal(d.cpu().numpy(), ed)


@pytest.mark.parametrize("name", ["G.mtx", "G_cc2.mtx", "rome99.gr"])
def test_fixture_roundtrip_and_bfs(hn, name):
    print(f"case fixture {name}", flush=True)
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)[name]
    off, adj = load_fixture(name)
    h = hn.Hornet.from_csr(off, adj)
    assert h.nV() == g["nv"] and h.nE() == g["ne"]
    s, d = export(h)
    es, ed = coo(off, adj)
    # export is sorted by (src, dst); the input may hold arcs in file order
    order = np.lexsort((ed, es))
    assert np.array_equal(s, es[order]) and np.array_equal(d, ed[order])
    assert h.max_degree_id() == g["max_degree_id"]
    bfs = hn.BfsTopDown(h)
    for tag in ("bfs_maxdeg", "bfs_0"):
        bfs.set_parameters(g[tag]["source"])
        bfs.run()
        assert bfs.distances().cpu().numpy().tolist() == g[tag]["dist"], tag


def test_device_coo_build_matches_oracle(hn):
    print("case coo build", flush=True)
    import torch
    n = 1 << 13
    es, ed = O.uniform_pairs(n, 40000, seed=21)
    es[:100] = ed[:100]                       # some self loops
    h = hn.Hornet.from_coo_device(n, torch.from_numpy(es).cuda(), torch.from_numpy(ed).cuda())
    off, adj = O.coo_to_csr(n, es, ed)
    s, d = export(h)
    xs, xd = coo(off, adj)
    assert h.nE() == adj.size
    assert np.array_equal(s, xs) and np.array_equal(d, xd)


@pytest.fixture(scope="module")
def rmat16(hn):
    off, adj = O.rmat_graph(16, seed=1)
    h = hn.Hornet.rmat(16, seed=1)
    return off, adj, h


def test_rmat_graph_matches_oracle(rmat16):
    print("case rmat16 graph", flush=True)
    off, adj, h = rmat16
    assert h.nV() == 65536 and h.nE() == adj.size == 1819292
    s, d = export(h)
    xs, xd = coo(off, adj)
    assert np.array_equal(s, xs) and np.array_equal(d, xd)
    deg = h.degrees().cpu().numpy()
    assert np.array_equal(deg, np.diff(off))
    assert h.max_degree_id() == O.max_degree_id(off)
    assert h.max_degree() == int(np.diff(off).max())


def test_bfs_rmat16_sources(hn, rmat16):
    off, adj, h = rmat16
    bfs = hn.BfsTopDown(h)
    srcs = [O.max_degree_id(off)] + O.pick_sources(off, 16, seed=3).tolist()
    for src in srcs:
        print(f"case bfs rmat16 src={src}", flush=True)
        exp, scanned = O.bfs(off, adj, src)
        bfs.set_parameters(src)
        bfs.run()
        assert np.array_equal(bfs.distances().cpu().numpy(), exp), src
        assert bfs.stats.edges_scanned == scanned
        assert bfs.stats.visited == int((exp != INF).sum())
        assert bfs.stats.levels == int(exp[exp != INF].max()) + 1


def test_bfs_isolated_source_and_tiny_graphs(hn, rmat16):
    print("case bfs edge cases", flush=True)
    off, adj, h = rmat16
    iso = int(np.flatnonzero(np.diff(off) == 0)[0])
    bfs = hn.BfsTopDown(h)
    bfs.set_parameters(iso)
    bfs.run()
    d = bfs.distances().cpu().numpy()
    assert d[iso] == 0 and np.all(np.delete(d, iso) == INF)
    # single vertex, no edges
    h1 = hn.Hornet.from_csr(np.array([0, 0]), np.zeros(0, np.int32))
    b1 = hn.BfsTopDown(h1)
    b1.set_parameters(0)
    b1.run()
    assert b1.distances().cpu().numpy().tolist() == [0]
    assert h1.getCOO()[0].numel() == 0


def test_bfs_level_loop_depths(hn, rmat16):
    """The level loop queues levels without a host round trip and reads the
    end flag after as many levels as the graph's previous traversal took:
    traversals of very different depth in a row (a 3000-vertex path from
    either end and its middle, R-MAT s16 between them), with and without
    kernel timing, each against the oracle."""
    off, adj, h = rmat16
    n = 3000
    s = np.arange(n - 1, dtype=np.int32)
    poff, padj = O.coo_to_csr(n, s, s + 1, symmetrize=True)
    hp = hn.Hornet.from_csr(poff, padj)
    bp = hn.BfsTopDown(hp)
    br = hn.BfsTopDown(h)
    src_r = O.max_degree_id(off)
    for timed in (0, 1):
        for graph, b, o, a, src in (("path", bp, poff, padj, 0), ("rmat16", br, off, adj, src_r),
                                    ("path", bp, poff, padj, n - 1), ("path", bp, poff, padj, n // 2),
                                    ("rmat16", br, off, adj, src_r)):
            print(f"case bfs levels {graph} src={src} timed={timed}", flush=True)
            exp, scanned = O.bfs(o, a, src)
            b.set_parameters(src)
            b.stats.time_kernels = timed
            b.run()
            assert np.array_equal(b.distances().cpu().numpy(), exp), (graph, src)
            assert b.stats.levels == int(exp[exp != INF].max()) + 1
            assert b.stats.edges_scanned == scanned
            assert b.stats.expand_launches == b.stats.levels
            assert (b.stats.expand_ns > 0) == bool(timed)
    hp.release()


def test_bfs_hubs_spanning_many_tiles(hn):
    """Frontiers whose vertices' adjacency spans many expansion tiles (a tile
    holds 2048 groups of 4 entries): two 100 000-leaf hubs joined through some
    of the first hub's leaves, searched from a hub, from a leaf and from the
    far side, symmetric and directed (where discovered vertices can have no
    out-entries), each against the oracle."""
    n = 200_000
    a_leaf = np.arange(1, 100_001, dtype=np.int32)
    b = 100_001
    b_leaf = np.arange(100_002, n, dtype=np.int32)
    bridge = a_leaf[::997]
    s = np.concatenate([np.zeros(a_leaf.size, np.int32), np.full(b_leaf.size, b, np.int32), bridge,
                        bridge[:5]])   # repeated bridge entries: a multigraph hub
    d = np.concatenate([a_leaf, b_leaf, np.full(bridge.size, b, np.int32), np.full(5, b, np.int32)])
    for sym in (True, False):
        off, adj = O.coo_to_csr(n, s, d, symmetrize=sym, dedup=False)
        h = hn.Hornet.from_csr(off, adj)
        bfs = hn.BfsTopDown(h)
        for src in (0, int(a_leaf[0]), b, int(b_leaf[-1]), int(bridge[3])):
            print(f"case bfs hubs sym={sym} src={src}", flush=True)
            exp, scanned = O.bfs(off, adj, src)
            bfs.set_parameters(src)
            bfs.run()
            assert np.array_equal(bfs.distances().cpu().numpy(), exp), (sym, src)
            assert bfs.stats.edges_scanned == scanned
            assert bfs.stats.visited == int((exp != INF).sum())
        h.release()


def test_invalid_inputs_raise(hn):
    print("case invalid inputs", flush=True)
    from hornet_amd import HornetError
    with pytest.raises(HornetError):
        hn.Hornet.from_csr(np.array([0, 2, 3]), np.array([1, 5, 0], np.int32))  # id out of range
    # rows in any order are accepted (HornetInit) and stored ascending
    hu = hn.Hornet.from_csr(np.array([0, 2, 3]), np.array([1, 0, 0], np.int32))
    us, ud = hu.getCOO()
    assert us.cpu().tolist() == [0, 0, 1] and ud.cpu().tolist() == [0, 1, 0]
    h = hn.Hornet.from_csr(np.array([0, 1, 2]), np.array([1, 0], np.int32))
    b = hn.BfsTopDown(h)
    b.set_parameters(7)
    with pytest.raises(HornetError):
        b.run()
