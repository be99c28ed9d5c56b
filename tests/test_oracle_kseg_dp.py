"""Pins the oracle's k-segmentation DP (segmentation.hpp:172-262) by brute
force: on a few 8-bit levels every way of cutting them into k contiguous
segments no longer than the extent e is enumerated, each scored with the
reference's integer cost (sum of squares minus the truncated square of the
sum over the segment length, segmentation.hpp:31-35), and the DP must give
the optimal cost and, among optimal cuts, the one its first-strict-minimum
backtrack picks: the smallest start of the last segment, then of the one
before it, and so on (segmentation.hpp:228-245, 252-257).
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle_lib as O


def seg_cost(q, a, b):
    """Integer cost of the segment q[a:b] (0-based, b exclusive)."""
    s = int(np.sum(q[a:b], dtype=np.int64))
    sq = int(np.sum(q[a:b].astype(np.int64) ** 2))
    return sq - (s * s) // (b - a)


def brute_force(q, e, k):
    n = len(q)
    best = None
    for cuts in itertools.combinations(range(1, n), k - 1):
        bp = (0,) + cuts + (n,)
        if any(bp[i + 1] - bp[i] > e for i in range(k)):
            continue
        c = sum(seg_cost(q, bp[i], bp[i + 1]) for i in range(k))
        key = (c, tuple(reversed(bp)))
        if best is None or key < best:
            best = key
    return best[0], list(reversed(best[1]))


def dp(q, e, k):
    L = O.lib()
    L.orc_kseg_dp.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    q = np.ascontiguousarray(q, dtype=np.uint8)
    cost = np.zeros(1, np.uint32)
    bp = np.zeros(k + 1, np.uint16)
    rc = L.orc_kseg_dp(O._p(q), len(q), e, k, O._p(cost), O._p(bp))
    assert rc == 0
    return int(cost[0]), bp.tolist()


@pytest.mark.parametrize("seed", range(8))
def test_kseg_dp_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(30):
        n = int(rng.integers(2, 11))
        if rng.random() < 0.5:
            q = rng.integers(0, 256, size=n)
        else:  # a few plateaus with noise: ties and clear optima both occur
            lv = rng.integers(0, 256, size=int(rng.integers(1, 4)))
            q = np.clip(np.repeat(lv, -(-n // len(lv)))[:n] + rng.integers(-2, 3, size=n), 0, 255)
        for k in range(1, min(n, 5) + 1):
            for e in sorted({1, 2, 3, n // 2 + 1, n, n + 3}):
                if e < 1 or k * e < n:
                    continue  # no cut of n values into k segments of <= e
                got = dp(q, e, k)
                exp = brute_force(q.astype(np.int64), e, k)
                assert got == exp, (q.tolist(), e, k, got, exp)
                checked += 1
    assert checked > 100


def test_kseg_dp_ties_pick_leftmost_last_start():
    # constant levels: every cut costs 0; the backtrack takes the smallest
    # start of the last segment allowed by e, then of each earlier one
    q = np.full(7, 40)
    assert dp(q, 7, 3) == (0, [0, 1, 2, 7])
    assert dp(q, 3, 3) == (0, [0, 1, 4, 7])


def test_kseg_dp_clean_steps():
    q = np.array([10] * 4 + [200] * 3 + [90] * 5)
    assert dp(q, 12, 3) == (0, [0, 4, 7, 12])
    c, bp = dp(q, 12, 2)
    assert c == brute_force(q, 12, 2)[0] and bp == brute_force(q, 12, 2)[1]
