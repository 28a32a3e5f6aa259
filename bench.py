"""Benchmark of the Hornet hot path on one MI355X.

Workload (one "step" = one pass of the hot path over one batch of input):
  R-MAT scale 24, edge factor 16 (a=.57 b=.19 c=.19), seed 24, int32 ids,
  symmetrized, self loops and duplicates removed, sorted adjacency, built on
  the device.  A step is BFS TopDown from 16 seeded non-isolated sources
  (seed 3) -- the BFS half of the headline metric.  GTEPS counts directed
  adjacency entries scanned (sum of degrees of reached vertices, m_C) per
  second; Graph500's convention (m_C / 2) is reported beside it.

Printed: ONE JSON line with the metric, the roofline of the dominant kernel
(bfs_expand_kernel, HIP events around every launch in the timed region) and
a CPU baseline (the oracle's sequential host BFS, restated from
xlib/src/Graph/BFS.cpp:54-76, on a bounded sample of the same sources).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BFS GTEPS + edge-updates/s at R-MAT s24; served p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
# L2 request ceiling: 8 XCDs x 16 L2 channels, one 128-byte request per
# channel per clock at the 2.4 GHz peak engine clock
L2_PEAK_GREQS = 8 * 16 * 2.4


def cpu_model() -> str:
    """The CPU as lscpu reports it: "Model name", with the family / model /
    stepping numbers when the model name alone is generic (virtual machines
    often report only the vendor's family name); else /proc/cpuinfo."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f.setdefault(k.strip(), v.strip())
        if "Model name" in f:
            ids = ", ".join(f"{k.lower()} {f[k]}" for k in ("CPU family", "Model", "Stepping") if k in f)
            return f["Model name"] + (f" ({ids})" if ids else "")
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


class one_core:
    """Pins the calling thread to one CPU for a timed host baseline (restored
    on exit)."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = min(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)
        return False


def median_time(fn, reps=3):
    """(median seconds, last result) over `reps` calls of fn()."""
    ts = []
    out = None
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2], out


def update_bytes(st):
    """Algorithmic HBM bytes of one batch (b submitted entries, b' applied,
    T touched, R relocated): insert 8b + 16b + 16|T| + 4 sum_T d_old + 4b' +
    8 sum_R d_old + 16|R|; erase the same with 8b' and sum_R d_new."""
    b = st["submitted"]
    per_applied = 4 if st["kind"] == 1 else 8
    return (24 * b + 16 * st["touched"] + 4 * st["touched_deg"] + per_applied * st["applied"]
            + 8 * st["relocated_deg"] + 16 * st["relocated"])


def load_traffic(name: str):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE for the
    same workload (committed summary), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def update_rounds(h, args, torch):
    """Dynamic updates on the same graph: per round r, insert `batch` uniform
    pairs (seed 4+r, ~1 % in-batch repeats, self loops dropped, symmetrized,
    removeBatchDuplicates + removeGraphDuplicates) and erase `batch` entries
    sampled uniformly from the current edge set (symmetrized,
    removeBatchDuplicates); recount triangles after every round."""
    from hornet_amd.graph import BatchUpdate, TriangleCounting2, uniform_pairs_device

    nv = h.nV()
    ins_t = era_t = 0.0
    ins_n = era_n = 0
    ins_added = era_removed = 0
    ins_bytes = era_bytes = 0
    moved = 0
    tc_t = 0.0
    tris = []
    tc = TriangleCounting2(h)
    tc.init()
    for r in range(args.rounds):
        us, ud = uniform_pairs_device(nv, args.batch, seed=4 + r)
        keep = us != ud
        bi = BatchUpdate(torch.cat([us[keep], ud[keep]]), torch.cat([ud[keep], us[keep]]))
        torch.cuda.synchronize()
        t = time.perf_counter()
        ins_added += h.insert(bi, True, True)
        torch.cuda.synchronize()
        ins_t += time.perf_counter() - t
        ins_n += bi.nE()
        st = h.update_stats()
        ins_bytes += update_bytes(st)
        moved += st["moved"]
        # deletions: positions uniform over the current directed entries
        src, dst = h.getCOO()
        pos, _ = uniform_pairs_device(src.numel(), args.batch, seed=100 + r)
        ps, pd = src[pos.long()], dst[pos.long()]
        be = BatchUpdate(torch.cat([ps, pd]), torch.cat([pd, ps]))
        del src, dst
        torch.cuda.synchronize()
        t = time.perf_counter()
        era_removed += h.erase(be, True)
        torch.cuda.synchronize()
        era_t += time.perf_counter() - t
        era_n += be.nE()
        st = h.update_stats()
        era_bytes += update_bytes(st)
        moved += st["moved"]
        t = time.perf_counter()
        tc.run_undirected()
        torch.cuda.synchronize()
        tc_t += time.perf_counter() - t
        tris.append(tc.countTriangles())
        print(f"# round {r}: ne={h.nE()} triangles={tris[-1]}", file=sys.stderr, flush=True)
    utr = load_traffic("update_traffic.json")
    updates = {
        "rounds": args.rounds, "pairs_per_round": args.batch,
        "edge_updates_per_s": round((ins_n + era_n) / (ins_t + era_t), 1),
        "insert_updates_per_s": round(ins_n / ins_t, 1),
        "erase_updates_per_s": round(era_n / era_t, 1),
        "insert_ms_avg": round(ins_t / args.rounds * 1e3, 3),
        "erase_ms_avg": round(era_t / args.rounds * 1e3, 3),
        "inserted": ins_added, "erased": era_removed, "final_ne": h.nE(),
        "unit": "directed edge-updates/s (submitted entries)",
        # whole insert / erase pipelines against HBM: the algorithmic bytes of
        # each batch (touched and relocated vertices measured per batch) over
        # its wall time
        "roofline": {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": round((ins_bytes + era_bytes) / (ins_t + era_t) / 1e9, 1),
            "frac": round((ins_bytes + era_bytes) / (ins_t + era_t) / 1e9 / HBM_PEAK_GBS, 4),
            "insert_GBps": round(ins_bytes / ins_t / 1e9, 1), "erase_GBps": round(era_bytes / era_t / 1e9, 1),
            "alg_bytes_per_insert": round(ins_bytes / max(args.rounds, 1)),
            "alg_bytes_per_erase": round(era_bytes / max(args.rounds, 1)),
            "entries_moved_per_batch": round(moved / max(2 * args.rounds, 1)),
            # HBM bytes per batch from the committed FETCH_SIZE / WRITE_SIZE
            # passes on the same workload (round 1 of these update rounds)
            "traffic": (utr or {}).get("bytes_per_batch"),
            "insert_traffic": (utr or {}).get("insert_bytes_per_batch"),
            "erase_traffic": (utr or {}).get("erase_bytes_per_batch"),
        },
    }
    tc_info = {"triangles_per_round": tris, "ms_avg": round(tc_t / args.rounds * 1e3, 3),
               "triangles_per_s": round(sum(tris) / tc_t, 1)}
    return updates, tc_info


def snapshot_timing(h, torch, reps=3):
    """hn_graph_snapshot of the resident graph: CSR export + rebuild into a
    tight arena (reads and writes every entry twice, plus the records)."""
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        s = h.snapshot()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
        s.release()
    ne, nv = h.nE(), h.nV()
    best = min(ts)
    moved = 16 * ne + 40 * nv    # 4 B read + 4 B written per entry, twice; records, offsets
    return {"ms": round(best * 1e3, 3), "ms_runs": [round(x * 1e3, 3) for x in ts],
            "effective_GBps": round(moved / best / 1e9, 1), "ne": ne}


def meta_rounds(h, args, torch):
    """The same insert / erase rounds on a copy of the graph that carries one
    4-byte field per edge (its position in the initial CSR); inserted entries
    carry their batch position.  Reports the rates beside the plain ones."""
    from hornet_amd.graph import BatchUpdate, Hornet, uniform_pairs_device

    off, adj = h.getCSR()
    w = torch.arange(adj.numel(), dtype=torch.int32, device=adj.device)
    torch.cuda.synchronize()
    t = time.perf_counter()
    hw = Hornet.from_device_csr(off, adj, meta=[w])
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t
    del off, adj, w
    nv = hw.nV()
    ins_t = era_t = 0.0
    ins_n = era_n = 0
    for r in range(args.meta_rounds):
        us, ud = uniform_pairs_device(nv, args.batch, seed=40 + r)
        keep = us != ud
        bs, bd = torch.cat([us[keep], ud[keep]]), torch.cat([ud[keep], us[keep]])
        bw = torch.arange(bs.numel(), dtype=torch.int32, device=bs.device).cpu().numpy()
        bi = BatchUpdate(bs, bd, meta=[bw])
        torch.cuda.synchronize()
        t = time.perf_counter()
        hw.insert(bi, True, True)
        torch.cuda.synchronize()
        ins_t += time.perf_counter() - t
        ins_n += bi.nE()
        src, dst = hw.getCOO()
        pos, _ = uniform_pairs_device(src.numel(), args.batch, seed=400 + r)
        ps, pd = src[pos.long()], dst[pos.long()]
        be = BatchUpdate(torch.cat([ps, pd]), torch.cat([pd, ps]))
        del src, dst
        torch.cuda.synchronize()
        t = time.perf_counter()
        hw.erase(be, True)
        torch.cuda.synchronize()
        era_t += time.perf_counter() - t
        era_n += be.nE()
    out = {"rounds": args.meta_rounds, "fields": "1 x int32", "build_s": round(t_build, 3),
           "insert_updates_per_s": round(ins_n / ins_t, 1), "erase_updates_per_s": round(era_n / era_t, 1),
           "insert_ms_avg": round(ins_t / args.meta_rounds * 1e3, 3),
           "erase_ms_avg": round(era_t / args.meta_rounds * 1e3, 3), "final_ne": hw.nE()}
    hw.release()
    return out


def batched_bfs(args, torch):
    """C4: 4,096 seeded sources in launches of 64 (bit-packed masks, captured
    level loop, autotuned load balancing) on R-MAT s23 and a Chung-Lu graph
    (8 M vertices, average degree 32, exponent 2.1), against one source per
    launch (hn_bfs_topdown) on the same GPU."""
    from hornet_amd import rng
    from hornet_amd.graph import BfsMulti, BfsTopDown, Hornet

    out = {}
    makers = [("rmat_s23", lambda: Hornet.rmat(23, seed=23)),
              ("chung_lu_8M_avg32_g2.1", lambda: Hornet.chung_lu(1 << 23, 32, 2.1, seed=8))]
    for name, make in makers:
        h = make()
        nv = h.nV()
        srcs = rng.pick_sources(h.degrees().cpu().numpy(), args.ms_sources, seed=3)
        ms = BfsMulti(h)
        T, A = ms.tune(srcs[:64])
        dist = torch.empty(64 * nv, dtype=torch.int32, device="cuda")
        ms.run(srcs[:64], out=dist)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        edges = levels = 0
        for i in range(0, srcs.size, 64):
            ms.run(srcs[i:i + 64], out=dist[: 64 * nv] if srcs[i:i + 64].size == 64 else None)
            edges += ms.stats.edges
            levels += ms.stats.levels
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        del dist
        bfs = BfsTopDown(h)
        plain_n = min(args.ms_plain, srcs.size)
        bfs.set_parameters(int(srcs[0]))
        bfs.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pe = 0
        for s in srcs[:plain_n]:
            bfs.set_parameters(int(s))
            bfs.run()
            pe += bfs.stats.edges_scanned
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        out[name] = {
            "nv": nv, "ne": h.nE(), "sources": int(srcs.size), "launches": int((srcs.size + 63) // 64),
            "batched_gteps": round(edges / tb / 1e9, 2), "batched_s": round(tb, 4),
            "plain_gteps": round(pe / tp / 1e9, 2), "plain_sources_timed": plain_n,
            "speedup_vs_plain": round((edges / tb) / (pe / tp), 2),
            "tuned_degree_threshold": T, "tuned_push_alpha": A, "tuned_lb_units_per_lane": ms.params()[2],
            "avg_levels_per_launch": round(levels / ((srcs.size + 63) // 64), 2),
        }
        print(f"# batched BFS {name}: {out[name]}", file=sys.stderr, flush=True)
        del bfs, ms
        h.release()
        torch.cuda.empty_cache()
    return out


def serve_trace(h, nreq, seed, torch):
    """C5 mix: 70 % single-source BFS, 20 % 2-hop counts, 10 % update batches
    of floor(10^U(3,6)) pairs (half inserts of uniform pairs, half erases of
    pairs sampled from the current edge set), symmetrized.  Sources are
    uniform over non-isolated vertices."""
    import numpy as np
    from hornet_amd.graph import Server, uniform_pairs_device

    rng = np.random.default_rng(seed)
    nv = h.nV()
    deg = h.degrees().cpu().numpy()
    nz = np.flatnonzero(deg > 0)
    kinds = rng.choice(4, nreq, p=[0.7, 0.2, 0.05, 0.05])
    src, dst = h.getCOO()
    trace = []
    for i, k in enumerate(kinds):
        if k == Server.BFS or k == Server.TWO_HOP:
            trace.append((int(k), int(nz[rng.integers(0, nz.size)])))
            continue
        m = int(10 ** rng.uniform(3, 6))
        if k == Server.INSERT:
            a, b = uniform_pairs_device(nv, m, seed=10_000 + i)
            a, b = a.cpu().numpy(), b.cpu().numpy()
            keep = a != b
            a, b = a[keep], b[keep]
        else:
            pos = torch.from_numpy(rng.integers(0, src.numel(), m)).cuda()
            a, b = src[pos].cpu().numpy(), dst[pos].cpu().numpy()
        trace.append((int(k), (np.concatenate([a, b]), np.concatenate([b, a]))))
    del src, dst
    return trace


def latency_stats(res, t_total):
    import numpy as np
    out = {"requests": int(res.size), "throughput_rps": round(res.size / t_total, 1)}
    lat = (res["t_done_ns"] - res["t_arrive_ns"]) / 1e6
    for name, kind in (("bfs", 0), ("two_hop", 1), ("update", -1)):
        m = res["kind"] >= 2 if kind < 0 else res["kind"] == kind
        if m.any():
            out[name] = {"n": int(m.sum()), "p50_ms": round(float(np.percentile(lat[m], 50)), 3),
                         "p99_ms": round(float(np.percentile(lat[m], 99)), 3)}
    q = res["kind"] <= 1
    if q.any():
        out["mean_batch"] = round(float(res["batch"][q].mean()), 2)
    out["failed"] = int((res["status"] != 0).sum())
    return out


def served(args, torch):
    """C5: R-MAT s25 resident; the trace is first played closed-loop (all at
    t=0) to find the saturating throughput, then with Poisson arrivals (seed 5)
    at 70 % of it; the plain form (one query per launch, no overlap) is run
    the same way on a prefix of the same trace."""
    import numpy as np
    from hornet_amd.graph import Hornet, Server

    def fresh():
        return Hornet.rmat(args.serve_scale, seed=args.serve_scale)

    h = fresh()
    trace = serve_trace(h, args.serve_requests, seed=5, torch=torch)
    out = {"graph": f"R-MAT s{args.serve_scale} ef16", "nv": h.nV(), "ne": h.nE()}
    h.release()
    results = {}
    for label, plain, n in (("served", False, len(trace)), ("plain", True, min(args.serve_plain, len(trace)))):
        # every replay starts from the same graph: the trace's updates change it
        sub = trace[:n]
        h = fresh()
        srv = Server(h, max_batch=64, slots=args.serve_slots, plain=plain)
        t = time.perf_counter()
        srv.replay(sub, np.zeros(n, np.int64))
        sat = n / (time.perf_counter() - t)
        srv.close()
        h.release()
        rate = 0.7 * sat
        arrivals = np.cumsum(np.random.default_rng(5).exponential(1e9 / rate, n)).astype(np.int64)
        h = fresh()
        srv = Server(h, max_batch=64, slots=args.serve_slots, plain=plain)
        t = time.perf_counter()
        res = srv.replay(sub, arrivals)
        el = time.perf_counter() - t
        out[label] = {"saturating_rps": round(sat, 1), "offered_rps": round(rate, 1), "slots": srv.slots(),
                      **latency_stats(res, el)}
        results[label] = res
        srv.close()
        h.release()
        torch.cuda.empty_cache()
        print(f"# serve {label}: {out[label]}", file=sys.stderr, flush=True)
    out["saturating_speedup_vs_plain"] = round(out["served"]["saturating_rps"] / out["plain"]["saturating_rps"], 2)
    n_plain = min(args.serve_plain, len(trace))
    out["plain_requests"] = n_plain
    # every request of the plain prefix: same graph version and digest when served
    p, q = results["plain"], results["served"][:n_plain]
    for f in ("kind", "status", "version", "reached", "dist_sum", "max_depth", "applied"):
        if not np.array_equal(p[f], q[f]):
            raise CheckFailed(f"served and plain results differ in {f}")
    out["served_equals_plain_on_prefix"] = True
    # latency at equal load: the served form offered the plain form's rate
    rate = out["plain"]["offered_rps"]
    arrivals = np.cumsum(np.random.default_rng(5).exponential(1e9 / rate, n_plain)).astype(np.int64)
    h = fresh()
    srv = Server(h, max_batch=64, slots=args.serve_slots)
    t = time.perf_counter()
    res = srv.replay(trace[:n_plain], arrivals)
    out["served_at_plain_rate"] = {"offered_rps": rate, **latency_stats(res, time.perf_counter() - t)}
    srv.close()
    h.release()
    torch.cuda.empty_cache()
    print(f"# serve served at the plain rate: {out['served_at_plain_rate']}", file=sys.stderr, flush=True)
    return out


def arena_info(h):
    """Edge-arena footprint of the resident graph: slots held by blocks and
    slots reserved, per stored entry (4 bytes a slot)."""
    used, cap = h.memory()
    ne = max(h.nE(), 1)
    return {"used_bytes_per_edge": round(4 * used / ne, 3), "capacity_bytes_per_edge": round(4 * cap / ne, 3),
            "used_GB": round(4 * used / 1e9, 3), "capacity_GB": round(4 * cap / 1e9, 3)}


def batch_sweep(h, args, torch):
    """Insert and erase pipelines at 2 M, 8 M and 32 M directed entries (1, 4
    and 16 M uniform pairs, symmetrized, removeBatchDuplicates +
    removeGraphDuplicates; the erase takes the same entries back) on a copy of
    the graph; each size runs twice and the second run is reported."""
    from hornet_amd.graph import BatchUpdate, uniform_pairs_device

    g = h.copy(slack=2.0)
    nv = g.nV()
    out = []
    for pairs in args.sweep:
        us, ud = uniform_pairs_device(nv, pairs, seed=7000 + pairs.bit_length())
        keep = us != ud
        b = BatchUpdate(torch.cat([us[keep], ud[keep]]), torch.cat([ud[keep], us[keep]]))
        del us, ud, keep
        for rep in range(2):
            torch.cuda.synchronize()
            t = time.perf_counter()
            g.insert(b, True, True)
            torch.cuda.synchronize()
            ti = time.perf_counter() - t
            si = g.update_stats()
            t = time.perf_counter()
            g.erase(b, True)
            torch.cuda.synchronize()
            te = time.perf_counter() - t
            se = g.update_stats()
        bi, be = update_bytes(si), update_bytes(se)
        out.append({"entries": b.nE(), "insert_ms": round(ti * 1e3, 3), "erase_ms": round(te * 1e3, 3),
                    "insert_updates_per_s": round(b.nE() / ti, 1), "erase_updates_per_s": round(b.nE() / te, 1),
                    "insert_hbm_frac": round(bi / ti / 1e9 / HBM_PEAK_GBS, 4),
                    "erase_hbm_frac": round(be / te / 1e9 / HBM_PEAK_GBS, 4),
                    "touched": si["touched"], "relocated": si["relocated"], "entries_moved": si["moved"]})
        print(f"# batch sweep: {out[-1]}", file=sys.stderr, flush=True)
        del b
    g.release()
    torch.cuda.empty_cache()
    return out


class CheckFailed(RuntimeError):
    pass


def cpu_baselines_and_checks(h, sources, updates, tc_info, args, torch):
    """Host baselines on one pinned core (median of 3): the oracle's sequential
    BFS (a restatement of xlib/src/Graph/BFS.cpp:54-76) from a bounded sample
    of the s24 sources, the host set-model insert of one 2 M-entry batch, and
    the reference's CPU configuration (R-MAT s16 seed 1: BFS from the
    max-degree vertex and the host triangle count of
    hornetsnest/test/TriangleTest2.cu:50-69).  The oracle outputs are then
    compared with the device's on the same inputs (outside every timed
    region); a mismatch stops the benchmark."""
    import numpy as np

    import oracle as O
    from hornet_amd.graph import BatchUpdate, BfsTopDown, Hornet, TriangleCounting2, uniform_pairs_device

    off, adj = h.getCSR()
    off = off.cpu().numpy()
    adj = adj.cpu().numpy()
    nv = h.nV()
    # s24 BFS: sources in order until the sample reaches about cpu_seconds
    bfs_times, bfs_m, used = [], 0, []
    with one_core() as pin:
        for s in sources:
            t, (dist, scanned) = median_time(lambda: O.bfs(off, adj, int(s)))
            bfs_times.append(t)
            bfs_m += scanned
            used.append((int(s), dist))
            if 3 * sum(bfs_times) >= args.cpu_seconds:
                break
        core = pin.core
    gbfs = BfsTopDown(h)
    for s, dist in used:
        gbfs.set_parameters(s)
        gbfs.run()
        if not np.array_equal(gbfs.distances().cpu().numpy(), dist):
            raise CheckFailed(f"BFS distances from source {s} differ from the oracle")
    cpu = {"value": round(bfs_m / sum(bfs_times) / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
           "sample": f"sequential host BFS (oracle restatement of xlib/src/Graph/BFS.cpp:54-76) from the first "
                     f"{len(used)} of the {len(sources)} sources on the s24 graph after the update rounds, "
                     f"median of 3 runs each, pinned to CPU {core}",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "checked_against_device": True}
    if updates is not None:
        us, ud = uniform_pairs_device(nv, args.batch, seed=1000)
        us, ud = us.cpu().numpy(), ud.cpu().numpy()
        keep = us != ud
        bs = np.concatenate([us[keep], ud[keep]])
        bd = np.concatenate([ud[keep], us[keep]])
        with one_core():
            t, (o2, a2, added) = median_time(lambda: O.insert(off, adj, bs, bd, True, True))
        g = h.copy(slack=1.2)
        got = g.insert(BatchUpdate(bs, bd), True, True)
        go, ga = g.getCSR()
        same = got == added and np.array_equal(go.cpu().numpy(), o2) and np.array_equal(ga.cpu().numpy(), a2)
        g.release()
        del go, ga
        if not same:
            raise CheckFailed("the device insert differs from the host set model")
        updates["cpu_baseline"] = {
            "value": round(bs.size / t, 1), "unit": "edge-updates/s", "cores": 1, "kind": "port",
            "sample": f"host set-model insert (oracle restatement of BatchUpdate semantics) of {bs.size} directed "
                      f"entries into the s24 graph, median of 3, one pinned core; the device result of the same "
                      f"insert is equal"}
    # the reference's CPU configuration, in full
    o1, a1 = O.rmat_graph(16, seed=1)
    src1 = O.max_degree_id(o1)
    with one_core():
        tb, (d1, m1) = median_time(lambda: O.bfs(o1, a1, src1))
        tt, tri1 = median_time(lambda: O.triangles(o1, a1))
    g1 = Hornet.from_csr(o1, a1)
    b1 = BfsTopDown(g1)
    b1.set_parameters(src1)
    b1.run()
    t1 = TriangleCounting2(g1)
    t1.init()
    t1.run_undirected()
    if not (np.array_equal(b1.distances().cpu().numpy(), d1) and t1.countTriangles() == tri1):
        raise CheckFailed("R-MAT s16 BFS / triangle count differ from the oracle")
    g1.release()
    c1 = {"graph": "R-MAT s16 ef16 seed 1", "nv": int(o1.size - 1), "ne": int(a1.size), "bfs_source": int(src1),
          "bfs_ms": round(tb * 1e3, 3), "bfs_gteps": round(m1 / tb / 1e9, 4), "triangles": int(tri1),
          "tc_ms": round(tt * 1e3, 3), "triangles_per_s": round(tri1 / tt, 1), "cores": 1, "kind": "port",
          "timing": "median of 3, one pinned core", "checked_against_device": True}
    # the s22 configuration (16 seeded sources + a triangle count) on all host
    # cores with the oracle's parallel checker, labelled as such
    T = O.checker_threads()
    o2, a2 = O.rmat_graph(22, seed=2, threads=T)
    srcs2 = O.pick_sources(o2, 16, seed=3)
    tb2, exp2 = median_time(lambda: O.bfs_many(o2, a2, srcs2, threads=T))
    tt2, tri2 = median_time(lambda: O.triangles(o2, a2, threads=T))
    g2 = Hornet.rmat(22, seed=2)
    b2 = BfsTopDown(g2)
    for s, (d, _) in zip(srcs2, exp2):
        b2.set_parameters(int(s))
        b2.run()
        if not np.array_equal(b2.distances().cpu().numpy(), d):
            raise CheckFailed(f"R-MAT s22 BFS from {s} differs from the oracle")
    t2 = TriangleCounting2(g2)
    t2.init()
    t2.run_undirected()
    if t2.countTriangles() != tri2:
        raise CheckFailed("R-MAT s22 triangle count differs from the oracle")
    del b2, t2
    g2.release()
    m2 = sum(m for _, m in exp2)
    c2 = {"graph": "R-MAT s22 ef16 seed 2", "nv": int(o2.size - 1), "ne": int(a2.size), "sources": 16,
          "bfs_s": round(tb2, 3), "bfs_gteps": round(m2 / tb2 / 1e9, 4), "triangles": int(tri2),
          "tc_s": round(tt2, 3), "triangles_per_s": round(tri2 / tt2, 1), "cores": T, "kind": "port",
          "timing": f"median of 3, the oracle's parallel checker on {T} host threads (an all-core variant of the "
                    f"sequential host path, not the one-core baseline)", "checked_against_device": True}
    c1["all_cores_s22"] = c2
    if tc_info is not None:
        tc_info["cpu_baseline"] = {"value": c1["triangles_per_s"], "unit": "triangles/s", "cores": 1, "kind": "port",
                                   "sample": "host triangle count (oracle restatement of "
                                             "hornetsnest/test/TriangleTest2.cu:50-69 on the DIRECTED_BY_DEGREE "
                                             "graph) of R-MAT s16 seed 1 in full, median of 3, one pinned core"}
    return cpu, c1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=24)
    ap.add_argument("--sources", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget for the CPU baseline sample (0 disables)")
    ap.add_argument("--rounds", type=int, default=10, help="dynamic-update rounds after the BFS steps")
    ap.add_argument("--batch", type=int, default=1 << 20, help="pairs inserted / deleted per round")
    ap.add_argument("--meta-rounds", type=int, default=2,
                    help="update rounds on a copy carrying one int32 field per edge (0 disables)")
    ap.add_argument("--snapshot", type=int, default=1, help="time hn_graph_snapshot of the graph (0 disables)")
    ap.add_argument("--ms-sources", type=int, default=4096, help="batched-BFS sources (0 disables)")
    ap.add_argument("--ms-plain", type=int, default=64, help="sources timed one per launch for the ratio")
    ap.add_argument("--serve-requests", type=int, default=2000, help="served-stream requests (0 disables)")
    ap.add_argument("--serve-plain", type=int, default=300, help="requests replayed in the plain form")
    ap.add_argument("--serve-scale", type=int, default=25)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--sweep", type=lambda x: [int(v) for v in x.split(",") if v], default=[1 << 20, 1 << 22, 1 << 24],
                    help="pairs per batch of the insert / erase size sweep (empty disables)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size reference "
                         "for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    from hornet_amd import rng
    from hornet_amd.graph import BfsTopDown, Hornet

    torch.cuda.set_device(0)
    if args.calibrate_copy_mib > 0:
        a = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    t_build = time.perf_counter()
    h = Hornet.rmat(args.scale, seed=args.seed, edge_factor=args.edge_factor)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    nv, ne = h.nV(), h.nE()
    deg = h.degrees().cpu().numpy()
    sources = rng.pick_sources(deg, args.sources, seed=3)
    print(f"# graph R-MAT s{args.scale} ef{args.edge_factor}: nv={nv} ne={ne} built in {t_build:.2f}s",
          file=sys.stderr, flush=True)

    bfs = BfsTopDown(h)

    def step(timed: bool):
        tot_m = tot_n = kns = launches = 0
        for s in sources:
            bfs.set_parameters(int(s))
            bfs.stats.time_kernels = 1 if timed else 0
            bfs.run()
            tot_m += bfs.stats.edges_scanned
            tot_n += bfs.stats.visited
            kns += bfs.stats.expand_ns
            launches += bfs.stats.expand_launches
        return tot_m, tot_n, kns, launches

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    agg_m = agg_n = agg_ns = agg_l = 0
    for _ in range(args.steps):
        m, n, kns, ln = step(True)
        agg_m += m
        agg_n += n
        agg_ns += kns
        agg_l += ln
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)

    gteps = agg_m / elapsed / 1e9
    # bfs_expand_kernel algorithmic bytes: 4 B per adjacency entry scanned,
    # 16 B vertex record + 4 B queue read per frontier vertex, 4 B distance +
    # 4 B queue write per discovered vertex -> 4 m_C + 28 n_C
    alg_bytes = 4 * agg_m + 28 * agg_n
    kernel_s = agg_ns * 1e-9
    achieved = alg_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    traffic = load_traffic("bfs_expand_traffic.json")
    l2req = load_traffic("bfs_expand_l2req.json")
    roofline = {
        "kernel": "bfs_expand_kernel",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": (traffic or {}).get("bytes_per_launch"),
        "alg_bytes_per_launch": round(alg_bytes / max(agg_l, 1)),
        "avg_launch_us": round(agg_ns / max(agg_l, 1) / 1e3, 2),
        "launches": agg_l,
        "entries_per_launch": round(agg_m / max(agg_l, 1)),
    }
    if l2req and l2req.get("read_req_per_entry"):
        # the bound the random visited probes meet: L1 -> L2 read requests
        # (TCP_TCC_READ_REQ per scanned entry, from the committed counter
        # pass on the same workload) over the live kernel time
        reqs = l2req["read_req_per_entry"] * agg_m
        roofline["l2_requests"] = {
            "bound": "l2 request rate", "unit": "G req/s", "peak": L2_PEAK_GREQS,
            "achieved": round(reqs / kernel_s / 1e9, 1) if kernel_s > 0 else 0.0,
            "frac": round(reqs / kernel_s / 1e9 / L2_PEAK_GREQS, 4) if kernel_s > 0 else 0.0,
            "req_per_entry": l2req["read_req_per_entry"],
        }

    updates = tc_info = ms_info = snap_info = meta_info = None
    if args.snapshot:
        snap_info = snapshot_timing(h, torch)
        print(f"# snapshot {snap_info['ms']} ms", file=sys.stderr, flush=True)
    if args.rounds > 0:
        updates, tc_info = update_rounds(h, args, torch)

    cpu = None
    c1 = None
    if args.cpu_seconds > 0:
        cpu, c1 = cpu_baselines_and_checks(h, sources, updates, tc_info, args, torch)

    arena = arena_info(h)
    sweep = batch_sweep(h, args, torch) if args.sweep else None

    if args.meta_rounds > 0:
        meta_info = meta_rounds(h, args, torch)
        print(f"# weighted updates: {meta_info}", file=sys.stderr, flush=True)

    if args.ms_sources > 0:
        del bfs
        h.release()
        torch.cuda.empty_cache()
        ms_info = batched_bfs(args, torch)
    serve_info = served(args, torch) if args.serve_requests > 0 else None

    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"BFS TopDown x{len(sources)} sources (seed 3) on R-MAT s{args.scale} "
                        f"ef{args.edge_factor} seed {args.seed} (a=.57,b=.19,c=.19), symmetrized, deduped",
            "scale": args.scale, "nv": nv, "ne": ne, "sources": len(sources),
            "load_balancing": "BinarySearch (wave64 LDS-staged)",
        },
        "gteps_graph500": round(gteps / 2, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_reference_config": c1,
        "arena": arena,
        "updates": updates,
        "update_batch_sweep": sweep,
        "triangle_counting": tc_info,
        "snapshot": snap_info,
        "updates_with_edge_meta": meta_info,
        "batched_bfs": ms_info,
        "served": serve_info,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
