"""hn_bfs_topdown_noalloc: the BFS TopDown entry point that allocates nothing
(its traversal scratch is a static device array) gives the same distances and
statistics as hn_bfs_topdown, through the C ABI, on two calls in a row -- the
second on a smaller graph, so stale scratch contents of the first call must
not leak into it -- and refuses a graph whose scratch exceeds the array."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _run(G, lib, name, h, src):
    import torch
    dist = torch.full((h.nV(),), -7, dtype=torch.int32, device="cuda")
    st = G._lib.HnBfsStats()
    rc = getattr(lib, name)(h.handle, int(src), ctypes.c_void_p(dist.data_ptr()), ctypes.byref(st),
                            G._stream())
    assert rc == 0, (name, rc, lib.hn_last_error())
    torch.cuda.synchronize()
    return dist.cpu().numpy(), (st.levels, st.visited, st.edges_scanned)


def test_noalloc_bfs_matches_allocating_entry_point():
    import hornet_amd.graph as G
    lib = G._lib.load()
    limit = ctypes.c_int64()
    assert lib.hn_bfs_topdown_noalloc_limit(ctypes.byref(limit)) == 0
    assert limit.value == 64 << 20
    # first a larger graph, then a smaller one (the second call sees the
    # first call's scratch contents)
    for scale, seed in ((16, 1), (12, 3)):
        print(f"no-alloc BFS: R-MAT s{scale} seed {seed}", flush=True)
        h = G.Hornet.rmat(scale, seed=seed)
        off, adj = O.rmat_graph(scale, seed=seed)
        srcs = O.pick_sources(off, 3, seed=scale)
        for s in srcs:
            d_na, st_na = _run(G, lib, "hn_bfs_topdown_noalloc", h, s)
            d_al, st_al = _run(G, lib, "hn_bfs_topdown", h, s)
            assert np.array_equal(d_na, d_al), f"s{scale} source {s}: distances differ"
            assert st_na == st_al, f"s{scale} source {s}: statistics differ"
            exp, _ = O.bfs(off, adj, int(s))
            assert np.array_equal(d_na, exp), f"s{scale} source {s}: distances differ from the oracle"
        h.release()


def test_noalloc_bfs_refuses_graphs_beyond_its_scratch():
    import torch

    import hornet_amd.graph as G
    lib = G._lib.load()
    # 8 M vertices need about 105 MiB of traversal scratch, beyond the 64 MiB array
    nv = 1 << 23
    e = torch.tensor([0, 1], dtype=torch.int32, device="cuda")
    h = G.Hornet.from_coo_device(nv, e[:1], e[1:])
    dist = torch.empty(nv, dtype=torch.int32, device="cuda")
    rc = lib.hn_bfs_topdown_noalloc(h.handle, 0, ctypes.c_void_p(dist.data_ptr()), None, G._stream())
    assert rc == G._lib.HN_E_CAPACITY
    assert b"scratch" in lib.hn_last_error()
    # the allocating entry point still serves it
    d, _ = _run(G, lib, "hn_bfs_topdown", h, 0)
    assert d[0] == 0 and d[1] == 1 and (d[2:] == 2**31 - 1).all()
    h.release()
