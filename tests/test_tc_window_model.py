"""CPU model of the window intersection of triangle counting
(tc_window_kernel in hornet_amd/csrc/tc.hip).

A wave intersects the shorter list S with the longer list L by 64-entry
windows: each lane finds its S entry in the current L window (first index
with L[p] >= a, six halving steps), then the S window moves on when its last
entry is not above the L window's, else the L window does; once the L window
has moved past, S entries up to the old window's last entry count only above
it.  Windows past a list's end repeat its last entry.  The count must equal
the per-entry binary searches of the other classes (each S entry counts once
if L holds it), on simple lists and on multigraph lists with repeats.
"""
import bisect

import numpy as np
import pytest

W = 64


def window_count(sp, lp):
    sn, ln = len(sp), len(lp)
    if sn == 0 or ln == 0:
        return 0
    lane = np.arange(W)

    def win(x, k, n):
        return x[np.minimum(k + lane, n - 1)]

    i = j = 0
    a, b = win(sp, 0, sn), win(lp, 0, ln)
    lim = -1
    c = 0
    while i < sn and j < ln:
        p = np.zeros(W, np.int64)
        st = 32
        while st:
            p += np.where(b[p + st - 1] < a, st, 0)
            st >>= 1
        y = b[p & (W - 1)]
        c += int(np.sum((p < W) & (y == a) & (i + lane < sn) & (a > lim)))
        amax, bmax = a[W - 1], b[W - 1]
        if amax <= bmax:
            i += W
            a = win(sp, i, sn)
            lim = -1
        else:
            j += W
            lim = bmax
            b = win(lp, j, ln)
    return c


def bsearch_count(sp, lp):
    c = 0
    for x in sp:
        k = bisect.bisect_left(lp, x)
        c += k < len(lp) and lp[k] == x
    return c


@pytest.mark.parametrize("repeats", [False, True])
def test_window_intersection_equals_binary_searches(repeats):
    rng = np.random.default_rng(11 if repeats else 12)
    for t in range(300):
        sn = int(rng.integers(0, 400))
        ln = int(rng.integers(sn, 1500))
        u = int(rng.choice([20, 60, 700, 5000, 1 << 20]))
        if repeats:
            sp = np.sort(rng.integers(0, u, sn))
            lp = np.sort(rng.integers(0, u, ln))
        else:
            sp = np.sort(rng.choice(u, min(sn, u), replace=False))
            lp = np.sort(rng.choice(u, min(ln, u), replace=False))
        assert window_count(sp, lp) == bsearch_count(sp, lp), (t, sn, ln, u)


def test_window_model_is_not_vacuous():
    # dropping the rule that S entries count only above the old L window
    # double-counts an entry repeated across an L window boundary
    sp = np.array([5] * 3 + [9] * 70)
    lp = np.array([1] * 62 + [5, 5] + [5] * 5 + [9] * 3)
    exp = bsearch_count(sp, lp)
    assert window_count(sp, lp) == exp

    def window_count_no_lim(sp, lp):
        sn, ln = len(sp), len(lp)
        lane = np.arange(W)
        i = j = 0
        a = sp[np.minimum(lane, sn - 1)]
        b = lp[np.minimum(lane, ln - 1)]
        c = 0
        while i < sn and j < ln:
            p = np.searchsorted(b, a, side="left")
            y = b[np.minimum(p, W - 1)]
            c += int(np.sum((p < W) & (y == a) & (i + lane < sn)))
            if a[W - 1] <= b[W - 1]:
                i += W
                a = sp[np.minimum(i + lane, sn - 1)]
            else:
                j += W
                b = lp[np.minimum(j + lane, ln - 1)]
        return c

    assert window_count_no_lim(sp, lp) != exp
