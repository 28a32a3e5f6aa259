"""GPU runs on graphs read from files: GraphStd (host reader in
libhornet_hip.so) -> HornetInit -> Hornet -> BFS / TriangleCounting2, from
Python and from the C++ headers (the BFSTest.cu / TriangleTest2.cu shape),
checked against the oracle and the golden fixtures."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from hornet_amd.graph import (BfsTopDown, GraphStd, Hornet, TriangleCounting2, parsing_prop as PP,
                              structure_prop as S)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DRIVER = os.path.join(HERE, "cpp", "_build", "hornet_driver")


def gold(name):
    return os.path.join(GOLD, name)


@pytest.fixture(scope="module")
def golden():
    with open(gold("golden.json")) as f:
        return json.load(f)


def oracle_read(name, structure=0, parse=0):
    if name.endswith(".mtx"):
        return O.read_market_std(gold(name), structure, parse)
    if name.endswith(".bin"):
        return O.read_binary(gold(name))
    return O.read_dimacs9_std(gold(name), structure, parse)


@pytest.mark.parametrize("name,structure", [("G.mtx", 0), ("G.mtx", S.UNDIRECTED), ("G_cc2.mtx", 0),
                                            ("G_cc2.mtx", S.UNDIRECTED), ("G.bin", 0), ("rome99.gr", 0)])
def test_bfs_on_file_graph(name, structure):
    print("case bfs", name, structure, flush=True)
    g = GraphStd(structure).read(gold(name))
    exp_g = oracle_read(name, structure)
    h = Hornet(g.hornet_init())
    assert h.nV() == exp_g["n"] and h.nE() == exp_g["adj"].size
    root = g.max_out_degree_id()
    bfs = BfsTopDown(h)
    bfs.set_parameters(root)
    bfs.run()
    exp, _ = O.bfs(exp_g["off"], exp_g["adj"], root)
    assert np.array_equal(bfs.distances().cpu().numpy(), exp)


@pytest.mark.parametrize("name", ["G.mtx", "G_cc2.mtx", "rome99.gr"])
def test_triangles_on_file_graph(golden, name):
    print("case tc", name, flush=True)
    structure = S.UNDIRECTED if name.endswith(".mtx") else 0
    g = GraphStd(structure).read(gold(name), PP.DIRECTED_BY_DEGREE | PP.SORT)
    h = Hornet(g.hornet_init())
    tc = TriangleCounting2(h)
    tc.init()
    tc.run()
    if name.endswith(".mtx"):
        assert tc.countTriangles() == golden[name]["triangles"]
    # the count on the stored orientation equals the oracle's count on it
    assert tc.countTriangles() == O.host_tc(g.csr_out_offsets(), g.csr_out_edges())


def test_binary_round_trip_through_device(tmp_path):
    print("case bin round trip", flush=True)
    g = GraphStd().read(gold("rome99.gr"), PP.RANDOMIZE | PP.SORT, seed=3)
    p = str(tmp_path / "r.bin")
    g.writeBinary(p)
    r = GraphStd().read(p)
    h = Hornet(r.hornet_init())
    off, adj = h.getCSR()
    exp = O.read_binary(p)
    assert np.array_equal(off.cpu().numpy(), exp["off"])
    assert np.array_equal(adj.cpu().numpy(), exp["adj"])   # SORT: rows already ascending


def run_driver(*args):
    assert os.path.exists(DRIVER), "build with __graft_entry__.build()"
    print("case driver", args[0], os.path.basename(args[1]), flush=True)
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    return out.stdout


def field(text, key):
    import re
    return int(re.search(rf"\b{key}=(-?\d+)", text).group(1))


@pytest.mark.parametrize("name", ["rome99.gr", "G.mtx", "G.bin"])
def test_cpp_bfsfile(tmp_path, name):
    out = run_driver("bfsfile", gold(name), str(tmp_path / "d.bin"))
    exp_g = oracle_read(name)
    root = int(np.argmax(np.diff(exp_g["off"])))
    assert field(out, "root") == root and field(out, "validate") == 1
    exp, _ = O.bfs(exp_g["off"], exp_g["adj"], root)
    assert np.array_equal(np.fromfile(tmp_path / "d.bin", np.int32), exp)


@pytest.mark.parametrize("name", ["G.mtx", "G_cc2.mtx"])
def test_cpp_tcfile(golden, name):
    out = run_driver("tcfile", gold(name))
    assert field(out, "total") == golden[name]["triangles"]


def test_unsorted_rows_sorted_on_device(tmp_path):
    # a default read keeps file order inside rows (and repeated entries); the
    # graph stores every row ascending, so its CSR is the SORT read of the file
    print("case unsorted rows", flush=True)
    s, d = O.rmat_pairs(14, seed=9, edge_factor=8)
    p = str(tmp_path / "u.mtx")
    with open(p, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate pattern general\n{1 << 14} {1 << 14} {s.size}\n")
        f.write("".join(f"{a + 1} {b + 1}\n" for a, b in zip(s.tolist(), d.tolist())))
    g = GraphStd().read(p)
    assert not all(np.all(np.diff(g.csr_out_edges()[g.csr_out_offsets()[v]:g.csr_out_offsets()[v + 1]]) >= 0)
                   for v in range(g.nV()))
    h = Hornet(g.hornet_init())
    off, adj = h.getCSR()
    exp = O.read_market_std(p, 0, O.P_SORT)
    assert np.array_equal(off.cpu().numpy(), exp["off"])
    assert np.array_equal(adj.cpu().numpy(), exp["adj"])
