"""CPU model of the in-place update passes (hornet_amd/csrc/update.hip).

Both BatchUpdate rewrites move entries inside a vertex's block in one pass of
copy tiles that run in any order: erase moves survivors down
(erase_gaps_kernel / erase_defer_kernel / erase_copy_kernel), insert moves the
old entries from the first one that changes up (insert_defer_kernel /
insert_move_kernel, then insert_batch_kernel).  A tile reads all its sources
before writing; entries whose destination may still be unread by another
tile are deferred to a side buffer and placed afterwards
(place_side_kernel).  This model restates that index math -- run layout,
per-tile deferral counts, item sources and destinations -- on random rows
and checks, for tile sizes far smaller than the kernels' so that rows span
many tiles:
  * no first-pass write of one tile lands on a source of another tile, and
    no slot is written twice;
  * running the tiles in a random order (each reading before writing), then
    the side buffer, then the batch entries, leaves every row equal to the
    reference result (erase: the list without the deleted slots; insert: the
    sorted merge);
  * deferring one entry fewer than the rule asks breaks the first check
    (the check is not vacuous).
"""
import bisect

import numpy as np
import pytest


def check_tiles(items, n_tiles):
    """items[t] = [(w, src, dst, deferred)]: first-pass writes vs other tiles' sources."""
    srcs = [set(x[1] for x in its) for its in items]
    writes = [set(x[2] for x in its if not x[3]) for its in items]
    seen = set()
    for t in range(n_tiles):
        assert not (writes[t] & seen), "a slot written twice"
        seen |= writes[t]
        for u in range(n_tiles):
            if u != t and writes[t] & srcs[u]:
                return False
    return True


def run_tiles(arena, items, rng):
    side = []
    for t in rng.permutation(len(items)):
        vals = [arena[x[1]] for x in items[t]]
        for x, val in zip(items[t], vals):
            if x[3]:
                side.append((x[2], val))
            else:
                arena[x[2]] = val
    for d, val in side:
        arena[d] = val


def erase_case(rng, T, short=0):
    """Random rows in power-of-two blocks, deletions, some rows relocated."""
    nv = int(rng.integers(1, 30))
    degs = [int(rng.choice([0, 1, 2, 3, 5, 8, 17, 40, 100, 300])) for _ in range(nv)]
    caps = [1 << max(0, int(np.ceil(np.log2(max(d, 1))))) for d in degs]
    offs = list(np.cumsum([0] + caps[:-1]))
    end = sum(caps)
    arena = [-1] * (3 * end + 16)
    for v in range(nv):
        arena[offs[v]:offs[v] + degs[v]] = [1000 * v + i for i in range(degs[v])]
    snap = list(arena)
    nxt = end
    runs, expect = [], {}
    for v in range(nv):
        d = degs[v]
        if d == 0 or rng.random() < 0.3:
            continue
        k = int(rng.integers(1, d + 1)) if rng.random() < 0.3 else int(rng.integers(1, min(d, 6) + 1))
        dele = sorted(int(x) for x in rng.choice(d, k, replace=False))
        newd = d - k
        if newd == 0:
            continue
        inplace = rng.random() < 0.7
        new_off, p = (offs[v], dele[0]) if inplace else (nxt, 0)
        if not inplace:
            nxt += caps[v]
        keep = np.ones(d, bool)
        keep[dele] = False
        expect[v] = (new_off, [snap[offs[v] + i] for i in range(d) if keep[i]])
        # a vertex's runs are consecutive: the run before each deleted key, then its tail run
        vr, prev, shift = [], 0, 0
        for x in dele:
            start = max(prev, p)
            vr.append((max(0, x - start), offs[v] + start, new_off + start - shift, inplace))
            shift += 1
            prev = x + 1
        start = max(prev, p)
        vr.append((max(0, d - start), offs[v] + start, new_off + start - shift, inplace))
        tail = len(runs) + len(vr) - 1
        runs += [r + (tail,) for r in vr]
    if not runs:
        return None
    L = np.array([r[0] for r in runs])
    P = np.r_[0, np.cumsum(L)]
    W = int(P[-1])
    n_tiles = (W + T - 1) // T
    src = [r[1] for r in runs]
    dst = [r[2] for r in runs]
    inp = [r[3] for r in runs]
    rtail = [r[4] for r in runs]

    def run_of(w):
        return int(np.searchsorted(P, w, side="right") - 1)

    tcnt = [0] * n_tiles
    for t in range(1, n_tiles):            # erase_defer_kernel
        i0 = t * T
        r = run_of(i0)
        rp = run_of(i0 - 1)
        if inp[r] and rtail[r] == rtail[rp]:
            d0 = dst[r] - P[r] + i0
            sprev = src[rp] - P[rp] + i0 - 1
            if sprev >= d0:
                c = min(sprev - d0 + 1, int(P[rtail[r] + 1]) - i0, T)
                tcnt[t] = max(0, c - short)
    items = []
    for t in range(n_tiles):               # erase_copy_kernel
        w0 = t * T
        its = []
        for w in range(w0, min(W, w0 + T)):
            s = run_of(w)
            its.append((w, src[s] - P[s] + w, dst[s] - P[s] + w, w - w0 < tcnt[t]))
        items.append(its)
    return arena, items, n_tiles, expect


def insert_case(rng, T, short=0):
    """Random sorted rows and sorted batch segments (stored copies first on ties)."""
    nv = int(rng.integers(1, 30))
    rows = []
    off = 0
    for v in range(nv):
        d = int(rng.choice([0, 1, 2, 5, 9, 30, 70, 200]))
        A = sorted(int(x) for x in rng.integers(0, 60, d))
        nb = int(rng.choice([0, 1, 1, 2, 3, 8, 40])) if rng.random() < 0.8 else 0
        B = sorted(int(x) for x in rng.integers(0, 60, nb))
        cap = 1 << max(5, int(np.ceil(np.log2(max(d, 1)))))
        rows.append((A, B, off, cap))
        off += cap
    arena = [-1] * (3 * off + 64)
    for A, B, o, c in rows:
        arena[o:o + len(A)] = A
    snap = list(arena)
    nxt = off
    segs = []
    for A, B, o, c in rows:
        if not B:
            continue
        newd = len(A) + len(B)
        if newd <= c:   # in place; old entries up to the smallest batch value stay
            new_off, p = o, bisect.bisect_right(A, B[0])
        else:
            new_off, p = nxt, 0
            nxt += 1 << int(np.ceil(np.log2(newd)))
        segs.append(dict(A=A, B=B, old=o, new=new_off, p=p, inplace=newd <= c,
                         ub=[bisect.bisect_right(A, b) for b in B]))
    if not segs:
        return None
    aP = np.r_[0, np.cumsum([len(s["A"]) - s["p"] for s in segs])]
    W = int(aP[-1])
    n_tiles = (W + T - 1) // T

    def seg_of(w):
        return int(np.searchsorted(aP, w, side="right") - 1)

    tcnt = [0] * n_tiles
    for t in range(n_tiles - 1):           # insert_defer_kernel
        i1 = (t + 1) * T
        j = seg_of(i1)
        if aP[j] < i1 and segs[j]["inplace"]:
            s = segs[j]
            r_last = s["p"] + (i1 - 1 - aP[j])
            c = bisect.bisect_left(s["B"], snap[s["old"] + r_last])
            tcnt[t] = max(0, min(c, i1 - max(int(aP[j]), t * T)) - short)
    items = []
    for t in range(n_tiles):               # insert_move_kernel
        w0 = t * T
        wend = min(W, w0 + T)
        its = []
        for w in range(w0, wend):
            s = segs[seg_of(w)]
            r = s["p"] + (w - aP[seg_of(w)])
            val = snap[s["old"] + r]
            its.append((w, s["old"] + r, s["new"] + r + bisect.bisect_left(s["B"], val), w >= wend - tcnt[t]))
        items.append(its)
    return arena, items, n_tiles, segs


@pytest.mark.parametrize("T", [4, 16, 64])
def test_erase_model(T):
    rng = np.random.default_rng(T)
    deferred = 0
    for _ in range(400):
        case = erase_case(rng, T)
        if case is None:
            continue
        arena, items, n_tiles, expect = case
        assert check_tiles(items, n_tiles)
        deferred += sum(x[3] for its in items for x in its)
        run_tiles(arena, items, rng)
        for new_off, want in expect.values():
            assert arena[new_off:new_off + len(want)] == want
    assert deferred > 0


@pytest.mark.parametrize("T", [4, 16, 64])
def test_insert_model(T):
    rng = np.random.default_rng(100 + T)
    deferred = 0
    for _ in range(400):
        case = insert_case(rng, T)
        if case is None:
            continue
        arena, items, n_tiles, segs = case
        assert check_tiles(items, n_tiles)
        deferred += sum(x[3] for its in items for x in its)
        run_tiles(arena, items, rng)
        for s in segs:                     # insert_batch_kernel
            for q, b in enumerate(s["B"]):
                arena[s["new"] + q + s["ub"][q]] = b
        for s in segs:
            n = len(s["A"]) + len(s["B"])
            assert arena[s["new"]:s["new"] + n] == sorted(s["A"] + s["B"])
    assert deferred > 0


@pytest.mark.parametrize("make", [erase_case, insert_case])
def test_one_short_deferral_is_caught(make):
    rng = np.random.default_rng(7)
    caught = False
    for _ in range(400):
        case = make(rng, 16, short=1)
        if case is not None and not check_tiles(case[1], case[2]):
            caught = True
            break
    assert caught
