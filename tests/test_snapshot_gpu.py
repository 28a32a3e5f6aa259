"""Snapshots (hn_graph_snapshot / Hornet.snapshot -> HornetStatic) and
checkpoint / resume through the binary graph format, checked against the
oracle."""
import numpy as np
import pytest

import oracle as O
from hornet_amd.graph import (BatchUpdate, BfsMulti, BfsTopDown, GraphStd, Hornet, HornetError,
                              HornetStatic)

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


def csr(h):
    off, adj = h.getCSR()
    return off.cpu().numpy(), adj.cpu().numpy()


def test_snapshot_is_isolated_from_updates():
    print("case snapshot isolation", flush=True)
    off, adj = O.rmat_graph(12, seed=1)
    h = Hornet.rmat(12, seed=1)
    snap = h.snapshot()
    assert isinstance(snap, HornetStatic)
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 5000, seed=7)
    s2 = np.concatenate([bs, bd]).astype(np.int32)
    d2 = np.concatenate([bd, bs]).astype(np.int32)
    added = h.insert(BatchUpdate(s2, d2), True, True)
    o1, a1, exp_added = O.insert(off, adj, s2, d2, True, True)
    assert added == exp_added > 0
    so, sa = csr(snap)
    assert np.array_equal(so, off) and np.array_equal(sa, adj)
    lo, la = csr(h)
    assert np.array_equal(lo, o1) and np.array_equal(la, a1)
    # queries on the snapshot see the old version
    src = O.max_degree_id(off)
    b = BfsTopDown(snap)
    b.set_parameters(src)
    b.run()
    exp, _ = O.bfs(off, adj, src)
    assert np.array_equal(b.distances().cpu().numpy(), exp)
    sources = O.pick_sources(off, 8, seed=3)
    got = BfsMulti(snap).run(sources).cpu().numpy().reshape(len(sources), n)
    for i, s in enumerate(sources):
        assert np.array_equal(got[i], O.bfs(off, adj, int(s))[0])
    # erase on the live graph leaves the snapshot as it was
    h.erase(BatchUpdate(s2[:100], d2[:100]))
    so, sa = csr(snap)
    assert np.array_equal(so, off) and np.array_equal(sa, adj)
    with pytest.raises(HornetError):
        snap.insert(BatchUpdate(s2[:4], d2[:4]))
    with pytest.raises(HornetError):
        snap.erase(BatchUpdate(s2[:4], d2[:4]))


def test_snapshot_of_empty_and_tiny_graphs():
    print("case snapshot tiny", flush=True)
    h = Hornet.from_csr(np.array([0, 0, 0]), np.array([], np.int32))
    s = h.snapshot()
    assert s.nV() == 2 and s.nE() == 0
    h = Hornet.from_csr(np.array([0, 1, 2]), np.array([1, 0], np.int32))
    s = h.snapshot(slack=3.0)
    so, sa = csr(s)
    assert so.tolist() == [0, 1, 2] and sa.tolist() == [1, 0]


def test_checkpoint_resume_binary(tmp_path):
    print("case checkpoint resume", flush=True)
    off, adj = O.rmat_graph(13, seed=2)
    h = Hornet.rmat(13, seed=2)
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 4000, seed=9)
    h.insert(BatchUpdate(bs, bd), True, True)
    o1, a1, _ = O.insert(off, adj, bs, bd, True, True)
    lo, la = csr(h)
    p = str(tmp_path / "ckpt.bin")
    GraphStd.from_csr(lo, la).writeBinary(p, 8)
    r = Hornet(GraphStd().read(p).hornet_init())
    ro, ra = csr(r)
    assert np.array_equal(ro, o1) and np.array_equal(ra, a1)
    src = O.max_degree_id(o1)
    b = BfsTopDown(r)
    b.set_parameters(src)
    b.run()
    assert np.array_equal(b.distances().cpu().numpy(), O.bfs(o1, a1, src)[0])
