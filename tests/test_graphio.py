"""Graph files (no GPU): the oracle's GraphStd restatement pinned against the
reference's own files and independent constructions, then the library's
readers / writers (host code in libhornet_hip.so) checked against the oracle.

Pins of the oracle:
  * example/G.bin is the reference's binary dump of example/G.mtx read with the
    default structure: it fixes the default direction rule, the file-order
    placement inside each row and the binary layout;
  * G.mtx read UNDIRECTED equals golden.json's symmetrised, deduplicated CSR
    (itself checked against networkx in make_golden.py);
  * DIRECTED_BY_DEGREE equals oracle.directed_by_degree on the same input (the
    restatement pinned by the triangle counts), the in-CSR is the transpose,
    RM_SINGLETON keeps exactly the non-isolated vertices of a symmetric graph,
    and RANDOMIZE is a relabelling (mapping labels back gives the input).
"""
import os

import numpy as np
import pytest

import oracle as O
from hornet_amd import _lib
from hornet_amd.graph import GraphStd, HornetError, parsing_prop as PP, structure_prop as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FILES = ["G.mtx", "G_cc2.mtx", "G_weighted.mtx", "rome99.gr"]


def gold(name):
    return os.path.join(GOLD, name)


def oracle_read(name, structure=0, parse=0, seed=0):
    if name.endswith(".mtx"):
        return O.read_market_std(gold(name), structure, parse, seed)
    return O.read_dimacs9_std(gold(name), structure, parse, seed)


def edge_set(off, adj):
    s = np.repeat(np.arange(off.size - 1), np.diff(off))
    return set(zip(s.tolist(), adj.tolist()))


# ---------------------------------------------------------------- oracle pins
def test_oracle_default_read_matches_reference_binary():
    b = O.read_binary(gold("G.bin"))
    g = O.read_market_std(gold("G.mtx"))
    assert b["n"] == g["n"] == 9
    assert b["structure"] == g["structure"] == O.GS_DIRECTED
    assert np.array_equal(b["off"], g["off"]) and np.array_equal(b["adj"], g["adj"])
    # rows in file order, not sorted: vertex 2 lists 1, 8, 3 (lines "3 2", "3 9", "3 4")
    assert g["adj"][g["off"][2]:g["off"][3]].tolist() == [1, 8, 3]


def test_oracle_undirected_matches_golden(tmp_path):
    import json
    golden = json.load(open(gold("golden.json")))
    for name in ("G.mtx", "G_cc2.mtx"):
        g = O.read_market_std(gold(name), S.UNDIRECTED, PP.SORT)
        assert g["off"].tolist() == golden[name]["offsets"]
        assert g["adj"].tolist() == golden[name]["adj"]


def test_oracle_directed_by_degree_matches_orientation():
    for name in ("G.mtx", "G_cc2.mtx"):
        sym = O.read_market_std(gold(name), S.UNDIRECTED, PP.SORT)
        g = O.read_market_std(gold(name), S.UNDIRECTED, PP.SORT | PP.DIRECTED_BY_DEGREE)
        o, a = O.directed_by_degree(sym["off"], sym["adj"])
        assert np.array_equal(g["off"], o) and np.array_equal(g["adj"], a)
    sym = O.read_dimacs9_std(gold("rome99.gr"), 0, PP.SORT)
    g = O.read_dimacs9_std(gold("rome99.gr"), 0, PP.SORT | PP.DIRECTED_BY_DEGREE)
    o, a = O.directed_by_degree(sym["off"], sym["adj"])
    assert edge_set(g["off"], g["adj"]) == edge_set(o, a)


def test_oracle_ingoing_is_transpose():
    g = O.read_market_std(gold("G_cc2.mtx"), S.DIRECTED | S.ENABLE_INGOING)
    out = edge_set(g["off"], g["adj"])
    inn = edge_set(g["in_off"], g["in_adj"])
    assert inn == {(d, s) for s, d in out}


def test_oracle_rm_singleton_keeps_non_isolated():
    for name in ("G.mtx", "G_cc2.mtx"):
        full = O.read_market_std(gold(name), S.UNDIRECTED, PP.SORT)
        g = O.read_market_std(gold(name), S.UNDIRECTED, PP.SORT | PP.RM_SINGLETON)
        deg = np.diff(full["off"])
        kept = np.flatnonzero(deg > 0)
        assert g["n"] == kept.size
        back = {(int(kept[s]), int(kept[d])) for s, d in edge_set(g["off"], g["adj"])}
        assert back == edge_set(full["off"], full["adj"])


def test_oracle_randomize_is_relabelling():
    base = O.read_dimacs9_std(gold("rome99.gr"), 0, PP.SORT)
    g = O.read_dimacs9_std(gold("rome99.gr"), 0, PP.SORT | PP.RANDOMIZE, seed=11)
    perm = O.vertex_permutation(base["n"], 11)
    assert sorted(perm.tolist()) == list(range(base["n"]))
    inv = np.argsort(perm)
    back = {(int(inv[s]), int(inv[d])) for s, d in edge_set(g["off"], g["adj"])}
    assert back == edge_set(base["off"], base["adj"])


# -------------------------------------------------------- library vs oracle
STRUCTS = [0, S.DIRECTED, S.UNDIRECTED, S.DIRECTED | S.ENABLE_INGOING]
PARSES = [0, PP.SORT, PP.SORT | PP.DIRECTED_BY_DEGREE, PP.RM_SINGLETON, PP.SORT | PP.RM_SINGLETON,
          PP.RANDOMIZE, PP.RANDOMIZE | PP.SORT | PP.PRINT_INFO, PP.DIRECTED_BY_DEGREE]


def check_same(g: GraphStd, exp):
    assert g.nV() == exp["n"] and g.nE() == exp["adj"].size
    assert np.array_equal(g.csr_out_offsets(), exp["off"])
    assert np.array_equal(g.csr_out_edges(), exp["adj"])
    if exp["in_off"] is not None:
        assert np.array_equal(g.csr_in_offsets(), exp["in_off"])
        assert np.array_equal(g.csr_in_edges(), exp["in_adj"])


@pytest.mark.parametrize("name", FILES)
@pytest.mark.parametrize("structure", STRUCTS)
@pytest.mark.parametrize("parse", PARSES)
def test_reader_matches_oracle(name, structure, parse):
    try:
        exp = oracle_read(name, structure, parse, seed=5)
    except ValueError:
        with pytest.raises(HornetError) as e:
            GraphStd(structure).read(gold(name), parse, seed=5)
        assert e.value.status == _lib.HN_E_INVALID_ARG
        return
    g = GraphStd(structure).read(gold(name), parse, seed=5)
    check_same(g, exp)
    assert g.max_out_degree_id() == int(np.argmax(np.diff(exp["off"]))) if exp["n"] else True


def test_binary_reader_matches_oracle():
    g = GraphStd().read(gold("G.bin"))
    check_same(g, O.read_binary(gold("G.bin")))
    assert g.is_directed()


@pytest.mark.parametrize("offset_bytes", [4, 8])
@pytest.mark.parametrize("structure", [0, S.UNDIRECTED, S.DIRECTED | S.ENABLE_INGOING])
def test_binary_round_trip(tmp_path, offset_bytes, structure):
    g = GraphStd(structure).read(gold("rome99.gr" if structure != S.DIRECTED | S.ENABLE_INGOING
                                      else "G_cc2.mtx"), PP.SORT)
    p = str(tmp_path / "g.bin")
    g.writeBinary(p, offset_bytes)
    exp = O.read_binary(p)
    assert open(p, "rb").read(len("intlong" if offset_bytes == 8 else "intint")) == \
        (b"intlong" if offset_bytes == 8 else b"intint")
    r = GraphStd().read(p)
    check_same(r, exp)
    assert np.array_equal(r.csr_out_offsets(), g.csr_out_offsets())
    assert np.array_equal(r.csr_out_edges(), g.csr_out_edges())
    if structure & S.ENABLE_INGOING:
        assert np.array_equal(r.csr_in_offsets(), g.csr_in_offsets())
        assert np.array_equal(r.csr_in_edges(), g.csr_in_edges())


def test_writer_reproduces_reference_binary(tmp_path):
    p = str(tmp_path / "G.bin")
    GraphStd().read(gold("G.mtx")).writeBinary(p)
    assert open(p, "rb").read() == open(gold("G.bin"), "rb").read()


def test_market_round_trip(tmp_path):
    g = GraphStd().read(gold("rome99.gr"))
    p = str(tmp_path / "r.mtx")
    g.writeMarket(p)
    r = GraphStd(S.UNDIRECTED).read(p)   # a general file read UNDIRECTED: mirrored + deduplicated
    exp = O.read_market_std(p, S.UNDIRECTED)
    check_same(r, exp)
    r2 = GraphStd().read(p)
    assert np.array_equal(r2.csr_out_offsets(), g.csr_out_offsets())
    assert np.array_equal(r2.csr_out_edges(), g.csr_out_edges())


def write_mtx(path, n, s, d, symmetric=False, extra=""):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern " + ("symmetric" if symmetric else "general") + "\n")
        f.write("% a comment\n%\n")
        f.write(f"{n} {n} {len(s)}\n")
        for a, b in zip(s.tolist(), d.tolist()):
            f.write(f"{a + 1} {b + 1}\n")
        f.write(extra)


def test_large_file_parallel_parse(tmp_path):
    # ~1M entries (> 8 MiB of text): the body is parsed in several chunks
    s, d = O.rmat_pairs(16, seed=3, edge_factor=16)
    p = str(tmp_path / "big.mtx")
    write_mtx(p, 1 << 16, s, d)
    g = GraphStd().read(p, PP.SORT)
    check_same(g, O.read_market_std(p, 0, PP.SORT))
    g = GraphStd().read(p)
    check_same(g, O.read_market_std(p))


def test_symmetric_market_with_self_loops(tmp_path):
    s = np.array([0, 1, 2, 2, 3], np.int32)
    d = np.array([1, 2, 2, 0, 3], np.int32)
    p = str(tmp_path / "sym.mtx")
    write_mtx(p, 5, s, d, symmetric=True, extra="\n")
    for st, pa in [(0, 0), (0, PP.SORT), (S.UNDIRECTED, PP.RM_SINGLETON), (0, PP.DIRECTED_BY_DEGREE)]:
        check_same(GraphStd(st).read(p, pa), O.read_market_std(p, st, pa))
    g = GraphStd().read(p)
    assert g.nE() == 5 + 3          # self loops are not mirrored
    assert g.is_undirected()


def test_trailing_entries_beyond_header_ignored(tmp_path):
    p = str(tmp_path / "t.mtx")
    s, d = np.array([0, 1], np.int32), np.array([1, 2], np.int32)
    write_mtx(p, 3, s, d, extra="3 1\nnot an entry\n")
    g = GraphStd().read(p)
    assert g.nE() == 2 and g.csr_out_edges().tolist() == [1, 2]


@pytest.mark.parametrize("case", ["missing", "ext", "range", "short", "zero_id", "array", "bin_id",
                                  "bin_trunc", "gr_count", "both_dirs", "undirected_to_directed"])
def test_errors(tmp_path, case):
    p = str(tmp_path / "x.mtx")
    structure = 0
    if case == "missing":
        p = str(tmp_path / "none.mtx")
    elif case == "ext":
        p = str(tmp_path / "x.foo")
        open(p, "w").write("1 2\n")
    elif case == "range":
        write_mtx(p, 3, np.array([0]), np.array([3]))
    elif case == "short":
        open(p, "w").write("%%MatrixMarket matrix coordinate pattern general\n3 3 4\n1 2\n2 3\n")
    elif case == "zero_id":
        open(p, "w").write("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n0 2\n")
    elif case == "array":
        open(p, "w").write("%%MatrixMarket matrix array real general\n3 1\n1\n2\n3\n")
    elif case == "bin_id":
        p = str(tmp_path / "x.bin")
        open(p, "wb").write(b"longlong" + bytes(40))
    elif case == "bin_trunc":
        p = str(tmp_path / "x.bin")
        open(p, "wb").write(open(gold("G.bin"), "rb").read()[:-4])
    elif case == "gr_count":
        p = str(tmp_path / "x.gr")
        open(p, "w").write("c x\np sp 3 3\na 1 2 1\na 2 3 1\n")
    elif case == "both_dirs":
        structure = S.DIRECTED | S.UNDIRECTED
        p = gold("G.mtx")
    elif case == "undirected_to_directed":
        structure = S.DIRECTED
        p = gold("rome99.gr")
    with pytest.raises(HornetError) as e:
        GraphStd(structure).read(p)
    assert e.value.status == _lib.HN_E_INVALID_ARG
    assert _lib.load().hn_last_error()


def test_hornet_init_shape():
    g = GraphStd(S.UNDIRECTED).read(gold("G.mtx"), PP.SORT)
    init = g.hornet_init()
    assert init.nV == 9 and init.nE == 22
    assert init.offsets.dtype == np.int64 and init.edges.dtype == np.int32


def test_from_csr_binary_round_trip(tmp_path):
    exp = O.read_dimacs9_std(gold("rome99.gr"), 0, PP.SORT)
    g = GraphStd.from_csr(exp["off"], exp["adj"])
    assert g.is_undirected() and g.nV() == exp["n"]
    p = str(tmp_path / "c.bin")
    g.writeBinary(p)
    r = GraphStd().read(p)
    check_same(r, O.read_binary(p))
    assert np.array_equal(r.csr_out_offsets(), exp["off"]) and np.array_equal(r.csr_out_edges(), exp["adj"])


@pytest.mark.parametrize("off,adj", [([0, 2, 1], [0, 1]), ([0, 1, 3], [0, 1]), ([1, 1, 2], [0, 1]),
                                     ([0, 1, 2], [0, 2])])
def test_from_csr_rejects_bad_csr(off, adj):
    with pytest.raises(HornetError) as e:
        GraphStd.from_csr(np.array(off), np.array(adj, np.int32))
    assert e.value.status == _lib.HN_E_INVALID_ARG
