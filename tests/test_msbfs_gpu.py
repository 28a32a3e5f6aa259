"""GPU parity: batched multi-source BFS (bit-packed source masks) against the
oracle's single-source BFS, per source, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


@pytest.fixture(scope="module")
def g14(hn):
    off, adj = O.rmat_graph(14, seed=23)
    return off, adj, hn.Hornet.rmat(14, seed=23)


def expected(off, adj, srcs):
    return np.stack([O.bfs(off, adj, int(s))[0] for s in srcs])


def test_64_sources_rmat(hn, g14):
    print("case msbfs 64 sources", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, 64, seed=3)
    ms = hn.BfsMulti(h)
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))
    assert ms.stats.edges == sum(O.bfs(off, adj, int(s))[1] for s in srcs)
    assert ms.stats.push_levels >= 1 and ms.stats.pull_levels >= 1


@pytest.mark.parametrize("k", [1, 7, 33])
def test_k_sources_with_repeats_and_isolated(hn, g14, k):
    print(f"case msbfs k={k}", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, k, seed=10 + k)
    iso = int(np.flatnonzero(np.diff(off) == 0)[0])
    srcs[0] = iso
    if k > 2:
        srcs[2] = srcs[1]
    got = hn.BfsMulti(h).run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))


@pytest.mark.parametrize("T,alpha,ipt", [(1, 1, 4), (4096, 1, 2), (8, 1 << 20, 6), (64, 16, 2), (1 << 30, 4, 6)])
def test_lb_parameters_do_not_change_results(hn, g14, T, alpha, ipt):
    print(f"case msbfs T={T} alpha={alpha} units/lane={ipt}", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, 64, seed=5)
    ms = hn.BfsMulti(h)
    ms.set_params(T, alpha)
    ms.set_lb_tile(ipt)
    assert ms.params() == (T, alpha, ipt)
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))
    ms.set_lb_tile(4)


def test_tuner_picks_a_tile_and_keeps_results(hn, g14):
    print("case msbfs tuner", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, 64, seed=6)
    ms = hn.BfsMulti(h)
    T, A = ms.tune(srcs)
    t2, a2, ipt = ms.params()
    assert (t2, a2) == (T, A) and ipt in (2, 4, 6)
    assert np.array_equal(ms.run(srcs).cpu().numpy(), expected(off, adj, srcs))
    ms.set_params(32, 16)


def test_depth_limit_two_hop(hn, g14):
    print("case msbfs 2-hop", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, 20, seed=8)
    limits = np.array([2] * 10 + [0] * 10, np.int32)
    got = hn.BfsMulti(h).run(srcs, limits).cpu().numpy()
    full = expected(off, adj, srcs)
    for i, s in enumerate(srcs):
        if limits[i] == 2:
            exp = np.where(full[i] <= 2, full[i], INF)
            assert np.array_equal(got[i], exp)
            assert int((got[i] != INF).sum()) == O.two_hop(off, adj, int(s))
        else:
            assert np.array_equal(got[i], full[i])


def test_tuner_then_results(hn, g14):
    print("case msbfs tune", flush=True)
    off, adj, h = g14
    srcs = O.pick_sources(off, 64, seed=12)
    ms = hn.BfsMulti(h)
    T, a = ms.tune(srcs)
    assert T > 0 and a > 0
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))


def test_after_updates_sees_new_graph(hn):
    # degrees change with updates: the pull-side heavy list must follow
    print("case msbfs after updates", flush=True)
    off, adj = O.rmat_graph(12, seed=40)
    h = hn.Hornet.from_csr(off, adj)
    ms = hn.BfsMulti(h)
    ms.set_params(16, 16)
    srcs = O.pick_sources(off, 16, seed=1)
    ms.run(srcs)
    n = off.size - 1
    hub = 7
    bs = np.full(200, hub, np.int32)
    bd = np.random.default_rng(0).choice(n, 200, replace=False).astype(np.int32)
    bs, bd = np.concatenate([bs, bd]), np.concatenate([bd, bs])
    h.insert(hn.BatchUpdate(bs, bd), True, True)
    off2, adj2, _ = O.insert(off, adj, bs, bd, True, True)
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off2, adj2, srcs))


def test_chung_lu_generator_and_bfs(hn):
    print("case chung-lu", flush=True)
    n = 1 << 14
    s, d = hn.chung_lu_pairs_device(n, 16 * n, seed=9)
    es, ed = O.chung_lu_pairs(n, 16 * n, seed=9)
    assert np.array_equal(s.cpu().numpy(), es) and np.array_equal(d.cpu().numpy(), ed)
    h = hn.Hornet.chung_lu(n, 32, 2.1, seed=9)
    off, adj = O.coo_to_csr(n, es, ed)
    assert h.nE() == adj.size
    srcs = O.pick_sources(off, 64, seed=2)
    got = hn.BfsMulti(h).run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))


def test_directed_graph_runs_push_only(hn):
    # pulling reads out-adjacency as in-adjacency, which only a symmetric
    # graph allows: a directed graph must be traversed push-only
    print("case msbfs directed", flush=True)
    n = 1 << 12
    s, d = O.rmat_pairs(12, seed=31)
    off, adj = O.coo_to_csr(n, s, d, symmetrize=False)
    h = hn.Hornet.from_csr(off, adj)
    srcs = O.pick_sources(off, 48, seed=4)
    ms = hn.BfsMulti(h)
    ms.set_params(0, 1 << 20)   # would pull at every level if it could
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))
    assert ms.stats.pull_levels == 0


def test_symmetry_tracked_through_updates(hn):
    print("case msbfs symmetry after updates", flush=True)
    off, adj = O.rmat_graph(12, seed=41)
    h = hn.Hornet.from_csr(off, adj)      # symmetric, found by the full check
    srcs = O.pick_sources(off, 32, seed=6)
    ms = hn.BfsMulti(h)
    ms.set_params(0, 1 << 20)
    ms.run(srcs)
    assert ms.stats.pull_levels >= 1
    # a symmetric batch keeps pulling allowed
    n = off.size - 1
    rng = np.random.default_rng(3)
    a = rng.integers(0, n, 300).astype(np.int32)
    b = rng.integers(0, n, 300).astype(np.int32)
    a, b = a[a != b], b[a != b]
    h.insert(hn.BatchUpdate(np.concatenate([a, b]), np.concatenate([b, a])), True, True)
    off, adj, _ = O.insert(off, adj, np.concatenate([a, b]), np.concatenate([b, a]), True, True)
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))
    assert ms.stats.pull_levels >= 1
    # one direction only: the adjacency is no longer symmetric
    c = rng.integers(0, n, 200).astype(np.int32)
    e = rng.integers(0, n, 200).astype(np.int32)
    c, e = c[c != e], e[c != e]
    h.insert(hn.BatchUpdate(c, e), True, True)
    off, adj, _ = O.insert(off, adj, c, e, True, True)
    got = ms.run(srcs).cpu().numpy()
    assert np.array_equal(got, expected(off, adj, srcs))
    assert ms.stats.pull_levels == 0
