"""GPU parity at the configured sizes, bit-exact against the oracle.

C2  R-MAT s22 (seed 2): the device graph equals the oracle's CSR; BFS TopDown
    from the 16 seeded sources (seed 3), every distance vector and the BFS
    statistics; the TriangleCounting2 total.
C3  R-MAT s24 (seed 24): 10 rounds of +1 M uniform pairs (seed 4+r, symmetrized,
    removeBatchDuplicates + removeGraphDuplicates) and -1 M entries sampled
    uniformly from the current edge set (positions from seed 100+r,
    symmetrized, removeBatchDuplicates) -- the workload the benchmark times.
    After every round the exported CSR equals the host set model
    (hornet/test/HornetInsertTest.cu:52-75, HornetDeleteTest.cu:97-120 check
    the same multisets); the triangle count after the last round.
C4  one 64-source batched launch on R-MAT s23 (seed 23) and on the Chung-Lu
    graph (8 M vertices, average degree 32, exponent 2.1, seed 8), tuned as the
    benchmark tunes it: all 64 distance rows against the oracle BFS.
C5  a served prefix of the mixed stream on R-MAT s25: every request's digest
    and graph version equal between the served form (batched, overlapped) and
    the plain one-query-per-launch form, and the queries that saw the initial
    graph equal the oracle's.

The oracle work is split over threads (oracle.checker_threads()); each case
prints its stages so a slow step names itself.
"""
import time

import numpy as np
import pytest

import oracle as O

# the oracle side of C3 alone takes about two minutes on 16 host threads
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]
INF = 2**31 - 1


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


def say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def device_csr(h):
    off, adj = h.getCSR()
    return off.cpu().numpy(), adj.cpu().numpy()


def assert_same_csr(h, off, adj, what):
    go, ga = device_csr(h)
    assert go.size == off.size and ga.size == adj.size, f"{what}: sizes differ"
    assert np.array_equal(go, off), f"{what}: offsets differ"
    assert np.array_equal(ga, adj), f"{what}: adjacency differs"


def test_c2_rmat22_bfs16_and_triangles(hn):
    T = O.checker_threads()
    say("C2 R-MAT s22 seed 2: device build")
    h = hn.Hornet.rmat(22, seed=2)
    say(f"C2 oracle build ({T} threads)")
    off, adj = O.rmat_graph(22, seed=2, threads=T)
    assert_same_csr(h, off, adj, "C2 graph")
    srcs = O.pick_sources(off, 16, seed=3)
    say(f"C2 BFS from 16 sources, nE={adj.size}")
    bfs = hn.BfsTopDown(h)
    got = []
    for s in srcs:
        bfs.set_parameters(int(s))
        bfs.run()
        got.append((bfs.distances().cpu().numpy(), bfs.stats.visited, bfs.stats.edges_scanned))
    exp = O.bfs_many(off, adj, srcs, threads=T)
    for s, (gd, gv, gm), (ed, em) in zip(srcs, got, exp):
        assert np.array_equal(gd, ed), f"C2 BFS distances differ from source {s}"
        assert gv == int((ed != INF).sum()) and gm == em, f"C2 BFS statistics differ from source {s}"
    say("C2 triangle count")
    tc = hn.TriangleCounting2(h)
    tc.init()
    tc.run_undirected()
    assert tc.countTriangles() == O.triangles(off, adj, threads=T)
    h.release()


def test_c3_rmat24_ten_update_rounds(hn):
    T = O.checker_threads()
    nv = 1 << 24
    say("C3 R-MAT s24 seed 24: device build")
    h = hn.Hornet.rmat(24, seed=24)
    say(f"C3 oracle build ({T} threads)")
    off, adj = O.rmat_graph(24, seed=24, threads=T)
    assert_same_csr(h, off, adj, "C3 initial graph")
    for r in range(10):
        us, ud = O.uniform_pairs(nv, 1 << 20, seed=4 + r)
        keep = us != ud
        bs = np.concatenate([us[keep], ud[keep]])
        bd = np.concatenate([ud[keep], us[keep]])
        added = h.insert(hn.BatchUpdate(bs, bd), True, True)
        off, adj, e_added = O.insert(off, adj, bs, bd, True, True)
        assert added == e_added, f"C3 round {r}: inserted {added} != {e_added}"
        pos, _ = O.uniform_pairs(int(adj.size), 1 << 20, seed=100 + r)
        ps = (np.searchsorted(off, pos, side="right") - 1).astype(np.int32)
        pd = adj[pos]
        es = np.concatenate([ps, pd])
        ed = np.concatenate([pd, ps])
        removed = h.erase(hn.BatchUpdate(es, ed), True)
        off, adj, e_removed = O.erase(off, adj, es, ed, True)
        assert removed == e_removed, f"C3 round {r}: erased {removed} != {e_removed}"
        assert h.nE() == adj.size
        assert_same_csr(h, off, adj, f"C3 round {r}")
        say(f"C3 round {r}: +{added} -{removed} nE={adj.size} edge set equal")
    say("C3 triangle count after round 9")
    tc = hn.TriangleCounting2(h)
    tc.init()
    tc.run_undirected()
    assert tc.countTriangles() == O.triangles(off, adj, threads=T)
    h.release()


@pytest.mark.parametrize("graph", ["rmat_s23", "chung_lu_8M"])
def test_c4_batched_bfs_64_sources(hn, graph):
    T = O.checker_threads()
    say(f"C4 {graph}: device build")
    if graph == "rmat_s23":
        h = hn.Hornet.rmat(23, seed=23)
        off, adj = O.rmat_graph(23, seed=23, threads=T)
    else:
        h = hn.Hornet.chung_lu(1 << 23, 32, 2.1, seed=8)
        off, adj = O.chung_lu_graph(1 << 23, 32, 2.1, seed=8, threads=T)
    assert_same_csr(h, off, adj, f"C4 {graph} graph")
    srcs = O.pick_sources(off, 64, seed=3)
    ms = hn.BfsMulti(h)
    thr, alpha = ms.tune(srcs)
    say(f"C4 {graph}: tuned threshold={thr} alpha={alpha}; one 64-source launch")
    dist = ms.run(srcs).cpu().numpy()
    exp = O.bfs_many(off, adj, srcs, threads=T)
    edges = 0
    for i, (s, (ed, em)) in enumerate(zip(srcs, exp)):
        assert np.array_equal(dist[i], ed), f"C4 {graph}: distances differ for source {s} (row {i})"
        edges += em
    assert ms.stats.edges == edges
    h.release()


def _served_trace(off, adj, nreq, seed):
    """The benchmark's mix on the host: 70 % BFS, 20 % 2-hop, 5 % inserts of
    uniform pairs, 5 % erases of pairs sampled from the initial edge set, each
    batch floor(10^U(3,6)) pairs, symmetrized."""
    rng = np.random.default_rng(seed)
    n = off.size - 1
    nz = np.flatnonzero(np.diff(off) > 0)
    kinds = rng.choice(4, nreq, p=[0.7, 0.2, 0.05, 0.05])
    trace = []
    for i, k in enumerate(kinds):
        if k <= 1:
            trace.append((int(k), int(nz[rng.integers(0, nz.size)])))
            continue
        m = int(10 ** rng.uniform(3, 6))
        if k == 2:
            a, b = O.uniform_pairs(n, m, seed=10_000 + i)
            keep = a != b
            a, b = a[keep], b[keep]
        else:
            pos = rng.integers(0, adj.size, m)
            a = (np.searchsorted(off, pos, side="right") - 1).astype(np.int32)
            b = adj[pos]
        trace.append((int(k), (np.concatenate([a, b]), np.concatenate([b, a]))))
    return trace


def test_c5_served_prefix_matches_plain_s25(hn):
    T = O.checker_threads()
    scale = 25
    say("C5 R-MAT s25: device build")
    h = hn.Hornet.rmat(scale, seed=scale)
    off, adj = device_csr(h)
    h.release()
    trace = _served_trace(off, adj, 240, seed=5)
    results = {}
    for label, plain in (("plain", True), ("served", False)):
        h = hn.Hornet.rmat(scale, seed=scale)
        srv = hn.Server(h, max_batch=64, slots=2, plain=plain)
        # bursty arrivals (mean gap 0.5 ms) so the served form batches and overlaps
        arrivals = np.cumsum(np.random.default_rng(7).exponential(5e5, len(trace))).astype(np.int64)
        say(f"C5 replay {label}: {len(trace)} requests")
        results[label] = srv.replay(trace, arrivals if not plain else np.zeros(len(trace), np.int64))
        srv.close()
        h.release()
    p, s = results["plain"], results["served"]
    assert np.all(p["status"] == 0) and np.all(s["status"] == 0)
    assert s["batch"][s["kind"] <= 1].max() > 1, "the served form never batched"
    for f in ("kind", "version", "reached", "dist_sum", "max_depth", "applied"):
        assert np.array_equal(p[f], s[f]), f"C5: field {f} differs between served and plain"
    # queries that saw the initial graph, against the oracle
    say(f"C5 oracle build ({T} threads)")
    o_off, o_adj = O.rmat_graph(scale, seed=scale, threads=T)
    assert np.array_equal(o_off, off) and np.array_equal(o_adj, adj), "C5 graph differs from the oracle"
    del o_off, o_adj
    first = [i for i, (k, _) in enumerate(trace) if k <= 1 and s["version"][i] == 0][:8]
    bfs_ids = [i for i in first if trace[i][0] == 0]
    exp = O.bfs_many(off, adj, [trace[i][1] for i in bfs_ids], threads=T)
    for i, (d, _) in zip(bfs_ids, exp):
        r = d[d != INF].astype(np.int64)
        assert (int(s["reached"][i]), int(s["dist_sum"][i]), int(s["max_depth"][i])) == \
            (r.size, int(r.sum()), int(r.max())), f"C5 request {i} differs from the oracle"
    for i in first:
        if trace[i][0] == 1:
            assert int(s["reached"][i]) == O.two_hop(off, adj, trace[i][1]), f"C5 request {i} (2-hop)"
    assert len(first) > 0
