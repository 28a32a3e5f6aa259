"""GPU: the served mixed stream.  Every query result must equal the result of
the same query run alone on the graph version it arrived at (all updates
submitted before it, none after), whether it was batched with others,
overlapped with other launches, or served one per launch."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


@pytest.fixture(scope="module")
def hn():
    import hornet_amd.graph as G
    return G


def make_trace(off, adj, nreq, seed):
    """~70 % BFS, 20 % 2-hop, 10 % updates (half inserts, half erases of
    symmetric pairs), with Poisson arrivals of mean 30 us (bursty enough to
    batch and overlap)."""
    rng = np.random.default_rng(seed)
    n = off.size - 1
    deg = np.diff(off)
    nz = np.flatnonzero(deg > 0)
    s0 = np.repeat(np.arange(n), deg).astype(np.int32)
    trace = []
    for i in range(nreq):
        u = rng.random()
        if u < 0.7:
            trace.append((0, int(rng.choice(nz))))
        elif u < 0.9:
            trace.append((1, int(rng.choice(nz))))
        else:
            m = int(rng.integers(5, 400))
            if rng.random() < 0.5:
                a = rng.integers(0, n, m).astype(np.int32)
                b = rng.integers(0, n, m).astype(np.int32)
                k = a != b
                a, b = a[k], b[k]
                trace.append((2, (np.concatenate([a, b]), np.concatenate([b, a]))))
            else:
                p = rng.integers(0, adj.size, m)
                a, b = s0[p], adj[p]
                trace.append((3, (np.concatenate([a, b]), np.concatenate([b, a]))))
    arrivals = np.cumsum(rng.exponential(30_000, nreq)).astype(np.int64)
    return trace, arrivals


def expected_results(off, adj, trace):
    """Replay in arrival order on the host set model; digest each query."""
    out = []
    version = 0
    for kind, payload in trace:
        if kind == 0:
            d, _ = O.bfs(off, adj, payload)
            r = d[d != INF]
            out.append(("bfs", version, int(r.size), int(r.astype(np.int64).sum()), int(r.max())))
        elif kind == 1:
            out.append(("2hop", version, O.two_hop(off, adj, payload)))
        elif kind == 2:
            off, adj, _ = O.insert(off, adj, payload[0], payload[1], True, True)
            version += 1
            out.append(("upd", version))
        else:
            off, adj, _ = O.erase(off, adj, payload[0], payload[1], True)
            version += 1
            out.append(("upd", version))
    return out


def check(res, exp):
    assert np.all(res["status"] == 0)
    for row, e in zip(res, exp):
        if e[0] == "bfs":
            assert (int(row["version"]), int(row["reached"]), int(row["dist_sum"]), int(row["max_depth"])) == e[1:]
        elif e[0] == "2hop":
            assert (int(row["version"]), int(row["reached"])) == e[1:]
        else:
            assert int(row["version"]) == e[1]
    assert np.all(res["t_done_ns"] >= res["t_start_ns"])
    assert np.all(res["t_start_ns"] >= res["t_arrive_ns"])


@pytest.mark.parametrize("plain,slots", [(False, 2), (False, 1), (True, 1)])
def test_served_stream_matches_alone(hn, plain, slots):
    print(f"case serve plain={plain} slots={slots}", flush=True)
    off, adj = O.rmat_graph(12, seed=17)
    h = hn.Hornet.from_csr(off, adj)
    trace, arrivals = make_trace(off, adj, 400, seed=5)
    srv = hn.Server(h, max_batch=64, slots=slots, plain=plain)
    res = srv.replay(trace, arrivals)
    srv.close()
    check(res, expected_results(off, adj, trace))
    if not plain:
        # queries were batched
        assert res["batch"][res["kind"] <= 1].max() > 1


def test_budget_admission(hn):
    print("case serve budget", flush=True)
    from hornet_amd import HornetError
    off, adj = O.rmat_graph(10, seed=2)
    h = hn.Hornet.from_csr(off, adj)
    n = off.size - 1
    one_slot = hn.Server.slot_bytes(h, 64)
    assert one_slot > 24 * n
    with pytest.raises(HornetError):
        hn.Server(h, max_batch=64, slots=4, mem_budget_bytes=one_slot // 2)
    srv = hn.Server(h, max_batch=64, slots=4, mem_budget_bytes=int(one_slot * 2.5))
    assert srv.slots() == 2
    # an update larger than what the slots leave is refused, the graph unchanged
    big = np.arange(20000, dtype=np.int32) % n
    res = srv.replay([(2, (big, big[::-1].copy())), (0, 0)], np.array([0, 1000], np.int64))
    assert res["status"][0] == 2 and res["version"][1] == 0
    srv.close()
    s, d = h.getCOO()
    assert s.numel() == adj.size


def test_one_directional_updates_fall_back(hn):
    # windows batch queries across updates only when every update is a
    # symmetric batch; one-directional ones are applied one by one and the
    # graph is then traversed push-only
    print("case serve one-directional updates", flush=True)
    off, adj = O.rmat_graph(11, seed=19)
    h = hn.Hornet.from_csr(off, adj)
    trace, arrivals = make_trace(off, adj, 240, seed=8)
    rng = np.random.default_rng(1)
    n = off.size - 1
    for i, (kind, payload) in enumerate(trace):
        if kind == 2 and rng.random() < 0.5:
            a = rng.integers(0, n, 50).astype(np.int32)
            b = rng.integers(0, n, 50).astype(np.int32)
            k = a != b
            trace[i] = (2, (a[k], b[k]))
    srv = hn.Server(h, max_batch=64, slots=2)
    res = srv.replay(trace, arrivals)
    srv.close()
    check(res, expected_results(off, adj, trace))


def test_window_spans_updates(hn):
    # a burst of queries and updates queued together is served by launches
    # that span updates, each query still at its own version
    print("case serve window burst", flush=True)
    off, adj = O.rmat_graph(12, seed=29)
    h = hn.Hornet.from_csr(off, adj)
    trace, _ = make_trace(off, adj, 300, seed=12)
    arrivals = np.zeros(len(trace), np.int64)   # everything queued at once
    srv = hn.Server(h, max_batch=64, slots=1)
    res = srv.replay(trace, arrivals)
    srv.close()
    check(res, expected_results(off, adj, trace))
    q = res["kind"] <= 1
    assert res["batch"][q].max() >= 16


@pytest.mark.parametrize("slots", [1, 2])
def test_window_burst_with_failed_update(hn, slots):
    # a burst spanning several windows (with two slots, a window's set-up runs
    # while the previous launch is still on the device), one of whose updates
    # names a vertex id out of range: that update fails as a whole and changes
    # nothing, and every other result is the one its version gives alone
    print(f"case serve window burst with a failed update slots={slots}", flush=True)
    off, adj = O.rmat_graph(12, seed=31)
    n = off.size - 1
    h = hn.Hornet.from_csr(off, adj)
    trace, _ = make_trace(off, adj, 400, seed=14)
    upd = [i for i, (k, _) in enumerate(trace) if k >= 2]
    bad = upd[len(upd) // 2]
    a, b = trace[bad][1]
    a = a.copy()
    a[0] = n   # one id past the last vertex
    trace[bad] = (trace[bad][0], (a, b))
    srv = hn.Server(h, max_batch=64, slots=slots)
    res = srv.replay(trace, np.zeros(len(trace), np.int64))
    srv.close()
    ok = [t for i, t in enumerate(trace) if i != bad]
    exp = expected_results(off, adj, ok)
    assert res["status"][bad] != 0
    assert int(res["version"][bad]) == upd.index(bad)   # the updates before it all applied
    keep = np.ones(len(trace), bool)
    keep[bad] = False
    check(res[keep], exp)
    q = res["kind"] <= 1
    assert res["batch"][q].max() >= 16


def test_window_entries_capped_by_budget(hn):
    # a budget whose update share holds windows to about 600 update entries:
    # windows span fewer updates, and every result still matches its version
    print("case serve windows capped by the update budget", flush=True)
    off, adj = O.rmat_graph(12, seed=37)
    h = hn.Hornet.from_csr(off, adj)
    one_slot = hn.Server.slot_bytes(h, 64)
    trace, _ = make_trace(off, adj, 300, seed=16)
    srv = hn.Server(h, max_batch=64, slots=2, mem_budget_bytes=2 * one_slot + 2 * (2 * 96 + 9) * 600)
    assert srv.slots() == 2
    res = srv.replay(trace, np.zeros(len(trace), np.int64))
    srv.close()
    check(res, expected_results(off, adj, trace))
    assert res["batch"][res["kind"] <= 1].max() > 1
