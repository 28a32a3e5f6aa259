"""GPU parity of the C++ drop-in headers: a hipcc-built driver that uses
hornet::gpu::Hornet, BatchUpdate, BfsTopDown, TriangleCounting2 and the
forAll* operators, checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "hornet_driver")
INF = 2**31 - 1


def write_graph(path, off, adj):
    with open(path, "wb") as f:
        np.array([off.size - 1, adj.size], np.int64).tofile(f)
        off.astype(np.int64).tofile(f)
        adj.astype(np.int32).tofile(f)


def run(*args):
    assert os.path.exists(DRIVER), "build with __graft_entry__.build()"
    print("case driver", " ".join(args[:1]), flush=True)
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    return out.stdout


def field(text, key):
    return int(re.search(rf"\b{key}=(-?\d+)", text).group(1))


@pytest.fixture(scope="module")
def graph(tmp_path_factory):
    off, adj = O.rmat_graph(12, seed=6)
    p = str(tmp_path_factory.mktemp("g") / "g.bin")
    write_graph(p, off, adj)
    return off, adj, p


def test_bfs_topdown_operator_path(graph, tmp_path):
    off, adj, p = graph
    src = O.max_degree_id(off)
    out = run("bfs", p, str(src), str(tmp_path / "d.bin"))
    assert field(out, "max_degree_id") == src
    assert field(out, "validate") == 1 and field(out, "native_equal") == 1
    got = np.fromfile(tmp_path / "d.bin", np.int32)
    exp, _ = O.bfs(off, adj, src)
    assert np.array_equal(got, exp)
    assert field(out, "levels") == int(exp[exp != INF].max()) + 1


def test_triangle_counting2_oriented(graph, tmp_path):
    off, adj, _ = graph
    do, da = O.directed_by_degree(off, adj)
    p = str(tmp_path / "o.bin")
    write_graph(p, do, da)
    out = run("tc", p)
    assert field(out, "total") == field(out, "sum") == O.triangles(off, adj)


def test_operators_checksums(graph):
    off, adj, p = graph
    out = run("ops", p)
    s = np.repeat(np.arange(off.size - 1), np.diff(off))
    deg = np.diff(off)
    for lb in ("binary_search", "vertex_based4", "vertex_based64"):
        line = next(x for x in out.splitlines() if x.startswith(f"ops {lb} "))
        assert field(line, "edges") == adj.size
        assert field(line, "dst_sum") == int(adj.sum())
        assert field(line, "src_sum") == int(s.sum())
    higher = int((deg[s] > deg[adj]).sum())
    for lb in ("pairs_binary_search", "pairs_vertex_based1"):
        line = next(x for x in out.splitlines() if x.startswith(f"ops {lb} "))
        assert field(line, "higher") == higher
    assert field(out, "degree_sum") == adj.size
    even = s % 2 == 0
    line = next(x for x in out.splitlines() if x.startswith("ops queue_even "))
    assert field(line, "edges") == int(even.sum())
    assert field(line, "dst_sum") == int(adj[even].sum())
    assert field(out, "id_sum") == (off.size - 1) * (off.size - 2) // 2


def test_batch_update_through_headers(graph, tmp_path):
    off, adj, p = graph
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 2000, seed=31)
    s0 = np.repeat(np.arange(n), np.diff(off)).astype(np.int32)
    bs = np.concatenate([bs, s0[:40]]).astype(np.int32)
    bd = np.concatenate([bd, adj[:40]]).astype(np.int32)
    bp = tmp_path / "b.bin"
    with open(bp, "wb") as f:
        np.array([bs.size], np.int64).tofile(f)
        bs.tofile(f)
        bd.tofile(f)
    out = run("update", p, str(bp), str(tmp_path / "coo.bin"))
    o1, a1, added = O.insert(off, adj, bs, bd, True, True)
    half = bs.size // 2
    o2, a2, removed = O.erase(o1, a1, bs[:half], bd[:half])
    assert field(out, "inserted") == added and field(out, "erased") == removed
    raw = np.fromfile(tmp_path / "coo.bin", np.int8)
    pos = 0
    for (oo, aa) in ((o1, a1), (o2, a2)):
        m = int(raw[pos:pos + 8].view(np.int64)[0])
        pos += 8
        src = raw[pos:pos + 4 * m].view(np.int32)
        pos += 4 * m
        dst = raw[pos:pos + 4 * m].view(np.int32)
        pos += 4 * m
        es = np.repeat(np.arange(oo.size - 1), np.diff(oo))
        assert np.array_equal(src, es) and np.array_equal(dst, aa)


def test_operator_surface_and_atomic_bfs(tmp_path):
    # oriented (DIRECTED_BY_DEGREE) R-MAT: forAllAdjUnions credits |N(u) ∩ N(v)|
    # to the smaller-degree endpoint, which is TriangleCounting2's per-vertex rule
    off, adj = O.rmat_graph(12, seed=6)
    do, da = O.directed_by_degree(off, adj)
    p = str(tmp_path / "o.bin")
    write_graph(p, do, da)
    src = O.max_degree_id(do)
    out = run("ops2", p, str(src), str(tmp_path / "r.bin"))
    n = do.size - 1
    raw = open(tmp_path / "r.bin", "rb").read()
    dist = np.frombuffer(raw, np.int32, n)
    exp, _ = O.bfs(do, da, src)
    assert np.array_equal(dist, exp)
    assert field(out, "validate") == 1
    reached = int((exp != INF).sum())
    assert field(out, "enqueued") == reached - 1           # each vertex enqueued once
    assert field(out, "levels") == int(exp[exp != INF].max()) + 1
    ne = da.size
    assert field(out, "id_sum") == ne * (ne - 1) // 2
    assert field(out, "sum") == sum(range(0, n, 3))
    deg = np.diff(do)
    pair_line = next(x for x in out.splitlines() if x.startswith("ops2 vertex_pairs"))
    assert field(pair_line, "sum") == int(sum(deg[v] * 3 + deg[v + 1] for v in range(0, n - 1, 2)))
    line = next(x for x in out.splitlines() if x.startswith("ops2 edges_default"))
    s = np.repeat(np.arange(n), deg)
    assert field(line, "edges") == ne and field(line, "dst_sum") == int(da.sum())
    assert field(line, "src_sum") == int(s.sum())
    per_exp = O.tc2_per_vertex(do, da)
    tri = O.triangles(off, adj)
    for k, tag in enumerate(("wf0", "wf1", "wf1000", "queue")):
        per = np.frombuffer(raw, np.uint64, n, 4 * n + 8 * n * k)
        assert np.array_equal(per, per_exp), tag
        line = next(x for x in out.splitlines() if x.startswith(f"ops2 adj_unions_{tag} "))
        assert field(line, "total") == tri
        if tag == "wf0":
            assert field(line, "search_calls") == 0 and field(line, "merge_calls") > 0
        if tag == "wf1000":
            assert field(line, "search_calls") > 0


def test_snapshot_and_checkpoint_through_headers(graph, tmp_path):
    off, adj, p = graph
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 3000, seed=41)
    bp = tmp_path / "b.bin"
    with open(bp, "wb") as f:
        np.array([bs.size], np.int64).tofile(f)
        bs.astype(np.int32).tofile(f)
        bd.astype(np.int32).tofile(f)
    out = run("snapshot", p, str(bp), str(tmp_path / "ckpt.bin"), str(tmp_path / "coo.bin"))
    o1, a1, added = O.insert(off, adj, bs, bd, True, True)
    assert field(out, "inserted") == added and added > 0
    assert field(out, "frozen_ne") == adj.size
    assert field(out, "live_ne") == field(out, "resumed_ne") == a1.size
    raw = np.fromfile(tmp_path / "coo.bin", np.int8)
    pos = 0
    for (oo, aa) in ((off, adj), (o1, a1)):    # frozen copy = before; resumed = after
        m = int(raw[pos:pos + 8].view(np.int64)[0])
        pos += 8
        src = raw[pos:pos + 4 * m].view(np.int32)
        pos += 4 * m
        dst = raw[pos:pos + 4 * m].view(np.int32)
        pos += 4 * m
        assert np.array_equal(src, np.repeat(np.arange(oo.size - 1), np.diff(oo)))
        assert np.array_equal(dst, aa)
    # the checkpoint is the reference's binary format ("intlong": int64 offsets)
    ck = O.read_binary(str(tmp_path / "ckpt.bin"))
    assert np.array_equal(ck["off"], o1) and np.array_equal(ck["adj"], a1)


def test_edge_meta_through_headers(graph, tmp_path):
    # TypeList<int, double> fields through HornetInit::insertEdgeData, a weighted
    # BatchUpdatePtr, a device operator reading edge.field<N>() and getCOO
    off, adj, p = graph
    n = off.size - 1
    bs, bd = O.uniform_pairs(n, 3000, seed=51)
    bp = tmp_path / "b.bin"
    with open(bp, "wb") as f:
        np.array([bs.size], np.int64).tofile(f)
        bs.astype(np.int32).tofile(f)
        bd.astype(np.int32).tofile(f)
    out = run("meta", p, str(bp), str(tmp_path / "coo.bin"))
    j = np.arange(adj.size)
    w0 = (j * 7 + 1).astype(np.int32)
    w1 = 0.5 * j
    i = np.arange(bs.size)
    o1, a1, (f0, f1), added = O.insert_meta(off, adj, [w0, w1], bs, bd,
                                           [(100000 + i).astype(np.int32), -1.0 * i])
    line = next(x for x in out.splitlines() if x.startswith("meta "))
    assert field(line, "inserted") == added and field(line, "ne") == a1.size
    assert field(line, "sum0") == int(f0.astype(np.int64).sum())
    assert field(line, "sum1") == int((2.0 * f1).astype(np.int64).sum())
    raw = open(tmp_path / "coo.bin", "rb").read()
    m = int(np.frombuffer(raw, np.int64, 1)[0])
    pos = 8
    src = np.frombuffer(raw, np.int32, m, pos); pos += 4 * m
    dst = np.frombuffer(raw, np.int32, m, pos); pos += 4 * m
    g0 = np.frombuffer(raw, np.int32, m, pos); pos += 4 * m
    g1 = np.frombuffer(raw, np.float64, m, pos)
    assert np.array_equal(src, np.repeat(np.arange(n), np.diff(o1))) and np.array_equal(dst, a1)
    assert np.array_equal(g0, f0) and np.array_equal(g1, f1)


def test_hornet_api_members(graph, tmp_path):
    """Hornet(), Hornet(nV) filled by an insert, getCSR (int32 and int64
    offsets), reset(HornetInit&), print(), HornetInit::insertEdgeData<N> and
    TriangleCounting2::validate (Hornet.cuh:101-129, HornetInit.cuh:78-79,
    HornetInitialize.i.cuh:152-204)."""
    off, adj, p = graph
    out = run("api", p, str(tmp_path / "api.bin"))
    assert "api default nv=0 ne=0" in out
    assert field(out, "inserted") == adj.size
    with open(tmp_path / "api.bin", "rb") as f:
        def csr():
            nv, ne = np.fromfile(f, np.int64, 2)
            return np.fromfile(f, np.int64, nv + 1), np.fromfile(f, np.int32, ne)
        g_off, g_adj = csr()
        r_off, r_adj = csr()
    assert np.array_equal(g_off, off) and np.array_equal(g_adj, adj)
    half = (off.size - 1) // 2
    exp_off = np.minimum(off, off[half])
    assert np.array_equal(r_off, exp_off) and np.array_equal(r_adj, adj[: off[half]])
    assert out.count("offsets32_equal=1") == 2
    tri = O.triangles(off, adj)
    assert field(out, "undirected") == tri and field(out, "validate_undirected") == 1
    assert field(out, "stored") == int(O.tc2_per_vertex(off, adj).sum()) and field(out, "validate_stored") == 1
    assert field(out, "nonzero0") == 0 and field(out, "sum1_x4") == adj.size * (adj.size - 1) // 2
    # print(): "v : degree | neighbours" for every vertex of the reset graph
    rows = [ln for ln in out.splitlines() if re.match(r"^\d+ : \d+ \|", ln)]
    assert len(rows) == off.size - 1
    for v in (0, 1, half - 1, half, off.size - 2):
        head, tail = rows[v].split("|")
        assert head.split() == [str(v), ":", str(exp_off[v + 1] - exp_off[v])]
        assert [int(x) for x in tail.split()] == adj[exp_off[v]:exp_off[v + 1]].tolist()
