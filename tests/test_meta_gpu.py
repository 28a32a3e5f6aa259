"""Edge meta-data on the device (hn_graph_create_csr_meta, hn_graph_insert_meta,
hn_graph_export_coo_meta): every field must follow its edge through build,
row sort, insert (in place and relocating, all dedup flags), erase and
snapshot, bit-exact against the oracle's meta-data model."""
import numpy as np
import pytest

import oracle as O
from hornet_amd import _lib
from hornet_amd.graph import BatchUpdate, Hornet, HornetError, HornetInit

pytestmark = pytest.mark.gpu


def fields_for(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32),
            rng.integers(0, 2**63, n, dtype=np.int64)]


def build(off, adj, fields):
    return Hornet(HornetInit(off.size - 1, adj.size, off, adj).insertEdgeData(*fields))


def check(h, off, adj, fields):
    s, d, m = h.getCOO(meta=True)
    es = np.repeat(np.arange(off.size - 1), np.diff(off))
    assert np.array_equal(s.cpu().numpy(), es)
    assert np.array_equal(d.cpu().numpy(), adj)
    assert len(m) == len(fields)
    for got, exp in zip(m, fields):
        assert np.array_equal(got.cpu().numpy(), np.asarray(exp).view(got.cpu().numpy().dtype))


@pytest.fixture(scope="module")
def base():
    off, adj = O.rmat_graph(12, seed=5)
    return off, adj, fields_for(adj.size, 1)


def test_build_and_export(base):
    print("case meta build", flush=True)
    off, adj, f = base
    h = build(off, adj, f)
    assert h.edge_meta_bytes() == [4, 8]
    v = h.device()
    assert v.nmeta == 2 and v.meta_bytes[0] == 4 and v.meta_bytes[1] == 8 and v.meta[0] and v.meta[1]
    check(h, off, adj, f)


@pytest.mark.parametrize("db,dg", [(False, False), (True, False), (False, True), (True, True)])
def test_insert_with_meta(base, db, dg):
    print("case meta insert", db, dg, flush=True)
    off, adj, f = base
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 6000, seed=3)
    s0 = np.repeat(np.arange(n), np.diff(off)).astype(np.int32)
    # stored pairs and in-batch repeats with different fields
    bs = np.concatenate([bs, s0[:300], bs[:200]]).astype(np.int32)
    bd = np.concatenate([bd, adj[:300], bd[:200]]).astype(np.int32)
    bf = fields_for(bs.size, 7)
    h = build(off, adj, f)
    added = h.insert(BatchUpdate(bs, bd, meta=bf), db, dg)
    o1, a1, f1, eadded = O.insert_meta(off, adj, f, bs, bd, bf, db, dg)
    assert added == eadded
    check(h, o1, a1, f1)


def test_insert_then_erase_chain(base):
    print("case meta chain", flush=True)
    off, adj, f = base
    n = off.size - 1
    h = build(off, adj, f)
    for r in range(3):
        bs, bd = O.uniform_pairs(n, 20000, seed=11 + r)   # hubs relocate, small rows grow in place
        bf = fields_for(bs.size, 20 + r)
        h.insert(BatchUpdate(bs, bd, meta=bf))
        off, adj, f, _ = O.insert_meta(off, adj, f, bs, bd, bf)
        check(h, off, adj, f)
        es = np.concatenate([bs[::3], np.repeat(np.arange(n), np.diff(off))[::5].astype(np.int32)])
        ed = np.concatenate([bd[::3], adj[::5]])
        removed = h.erase(BatchUpdate(es, ed))
        off, adj, f, erem = O.erase_meta(off, adj, f, es, ed)
        assert removed == erem
        check(h, off, adj, f)


def test_unsorted_rows_with_duplicates():
    print("case meta row sort", flush=True)
    rng = np.random.default_rng(4)
    n = 3000
    s = np.sort(rng.integers(0, n, 40000))
    d = rng.integers(0, 200, 40000).astype(np.int32)      # many repeated destinations per row
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, s + 1, 1)
    off = np.cumsum(off)
    f = fields_for(d.size, 9)
    h = build(off, d, f)
    eo, ea, ef = O.csr_meta(off, d, f)
    check(h, eo, ea, ef)


def test_plain_insert_gives_zero_fields_and_snapshot_keeps_fields(base):
    print("case meta zeros/snapshot", flush=True)
    off, adj, f = base
    n = off.size - 1
    h = build(off, adj, f)
    snap = h.snapshot()
    bs, bd = O.uniform_pairs(n, 500, seed=5)
    h.insert(BatchUpdate(bs, bd))
    zeros = [np.zeros(bs.size, np.int32), np.zeros(bs.size, np.int64)]
    o1, a1, f1, _ = O.insert_meta(off, adj, f, bs, bd, zeros)
    check(h, o1, a1, f1)
    check(snap, off, adj, f)


def test_meta_errors(base):
    print("case meta errors", flush=True)
    off, adj, f = base
    with pytest.raises(HornetError):
        HornetInit(off.size - 1, adj.size, off, adj).insertEdgeData(np.zeros(adj.size, np.int16))
    plain = Hornet.from_csr(off, adj)
    with pytest.raises(HornetError) as e:
        plain.insert(BatchUpdate([0], [1], meta=[np.zeros(1, np.int32)]))
    assert e.value.status == _lib.HN_E_INVALID_ARG


def test_erase_hub_rows_in_place_with_meta():
    """Fields follow their entries through in-place compaction of long rows
    (shifts of thousands of slots across many copy tiles)."""
    print("case meta erase hub rows", flush=True)
    from test_update_gpu import hub_erase_batch, hub_graph
    off, adj, hubs = hub_graph(4)
    f = fields_for(adj.size, 6)
    h = build(off, adj, f)
    for r in range(2):
        es, ed = hub_erase_batch(off, adj, hubs, 20 + r, (0.4, 0.08)[r])
        removed = h.erase(BatchUpdate(es, ed))
        off, adj, f, erem = O.erase_meta(off, adj, f, es, ed)
        assert removed == erem
        check(h, off, adj, f)


def test_insert_hub_rows_in_place_with_meta():
    """Fields follow their entries when long rows grow in place (old entries
    moving up across many copy tiles) and batch fields land beside them."""
    print("case meta insert hub rows", flush=True)
    from test_update_gpu import hub_graph, hub_insert_batch
    off, adj, hubs = hub_graph(6)
    f = fields_for(adj.size, 7)
    h = build(off, adj, f)
    bs, bd = hub_insert_batch(off, hubs, 31)
    bf = fields_for(bs.size, 8)
    h.insert(BatchUpdate(bs, bd, meta=bf))
    off, adj, f, _ = O.insert_meta(off, adj, f, bs, bd, bf)
    check(h, off, adj, f)
