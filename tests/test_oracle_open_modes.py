"""Pins the oracle's open-start, open-end and open DTW against textbook closed
forms computed here from plain cost matrices, independently of the oracle's
swath loop (dtw.hpp:39-259).

- Open end (dtw.hpp:196-200): the top row takes its left neighbour without
  the local cost, so its corner value is the minimum over the top row of
  the textbook global matrix G, and the distance is sqrtf(min_j G[-1, j]) / |Y|
  in float (dtw.hpp:250-255). With a swath width W < |X| the swath-left
  columns j = kW (k >= 1) also take the diagonal without the local cost
  (dtw.hpp:172-187); the closed form then changes only at those columns.
- Open start (dtw.hpp:131-156): row 0 costs nothing, so rows 1.. are
  subsequence DTW of Y[1:] with a free start column (Mueller, Fundamentals of
  Music Processing, 2015, sec. 7.2.3): S[0, j] = c(y1, x_j),
  S[i, 0] = sum_k<=i c(y_k, x_0); the distance is sqrtf(S[-1, -1]) / |Y|.
- Open (both): subsequence DTW with both ends free, sqrtf(min_j S[-1, j]),
  not divided (os xor oe is false).
- The open-end early stop (dtw.hpp:74-129) does not change the distance:
  the oracle run with the check switched off gives the same distance, and
  (without ties) the same traceback counts, while the stop does fire.

Each check is computed from the matrices alone; a dropped local cost, a
wrong divisor, a wrong swath-left column or a stop that fires when the
top-right value is not the column minimum fails one of them.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O


def _big(dtype):
    return np.finfo(dtype).max if np.issubdtype(dtype, np.floating) else np.iinfo(dtype).max


def textbook_global(y, x, dtype):
    """Textbook symmetric DTW with squared-difference cost; returns G and the
    reference step codes with the tie order DIAG, UP, RIGHT."""
    ly, lx = len(y), len(x)
    G = np.zeros((ly, lx), dtype=dtype)
    M = np.zeros((ly, lx), dtype=np.uint8)
    for i in range(ly):
        for j in range(lx):
            d = dtype(y[i] - x[j])
            dd = dtype(d * d)
            if i == 0:
                G[i, j] = dd if j == 0 else dtype(G[i, j - 1] + dd)
                M[i, j] = 255 if j == 0 else 2
                continue
            cands = [(dtype(G[i - 1, j - 1] + dd), 1)] if j > 0 else []
            cands.append((dtype(G[i - 1, j] + dd), 3))
            if j > 0:
                cands.append((dtype(G[i, j - 1] + dd), 2))
            v = min(c[0] for c in cands)
            G[i, j] = v
            M[i, j] = next(c[1] for c in cands if c[0] == v)
    return G, M


def subsequence_dtw(yy, x, dtype):
    """Subsequence DTW with a free start column (Mueller 2015, sec. 7.2.3)
    and its step codes (reference codes, DIAG, UP, RIGHT tie order; the free
    start row has no predecessor)."""
    n, lx = len(yy), len(x)
    S = np.zeros((n, lx), dtype=dtype)
    M = np.zeros((n, lx), dtype=np.uint8)
    for i in range(n):
        for j in range(lx):
            d = dtype(yy[i] - x[j])
            dd = dtype(d * d)
            if i == 0:
                S[i, j] = dd
                continue
            cands = [(dtype(S[i - 1, j - 1] + dd), 1)] if j > 0 else []
            cands.append((dtype(S[i - 1, j] + dd), 3))
            if j > 0:
                cands.append((dtype(S[i, j - 1] + dd), 2))
            v = min(c[0] for c in cands)
            S[i, j] = v
            M[i, j] = next(c[1] for c in cands if c[0] == v)
    return S, M


def top_row_entries(below, yt, x, dtype, w=1024):
    """Per column, the top row's value from below when it takes its left
    neighbour free: min(below[j] + c, below[j-1] + c), except that at the
    swath-left columns j = k*w (k >= 1) the diagonal comes without c."""
    out = []
    for j in range(len(x)):
        d = dtype(yt - x[j])
        dd = dtype(d * d)
        v = dtype(below[j] + dd)
        if j > 0:
            diag = below[j - 1] if j % w == 0 else dtype(below[j - 1] + dd)
            v = min(v, diag)
        out.append(v)
    return out


def top_row_minimum(below, yt, x, dtype, w):
    return min(top_row_entries(below, yt, x, dtype, w))


def f32_distance(cost, ly, divide, dtype):
    f = np.float32(np.sqrt(np.float32(cost)))
    if divide:
        f = np.float32(f / np.float32(ly))
    if np.issubdtype(dtype, np.integer):
        return dtype(np.trunc(f))
    return dtype(f)


def _pair(rng, dtype, ly, lx):
    if np.issubdtype(dtype, np.floating):
        return rng.normal(size=ly).astype(dtype), rng.normal(size=lx).astype(dtype)
    return rng.integers(-30, 30, size=ly).astype(dtype), rng.integers(-30, 30, size=lx).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("seed", range(6))
def test_open_end_is_min_over_textbook_top_row(dtype, seed):
    """|X| <= 1024: one swath, no swath-left column, no early stop."""
    rng = np.random.default_rng(100 + seed)
    ly, lx = int(rng.integers(2, 12)), int(rng.integers(1, 16))
    y, x = _pair(rng, dtype, ly, lx)
    G, Mg = textbook_global(y, x, dtype)
    m = G[-1].min()
    d, Mo, ab, _ = O.dtw_pair(y, x, os_=0, oe=1, moves=True)
    assert ab == -1
    assert d == f32_distance(m, ly, True, dtype), (d, m)
    # below the top row the matrix is the global one
    assert np.array_equal(Mo[:-1], Mg[:-1])
    # the walk leaves the top row at the last column whose entry from
    # below holds the minimum
    ent = top_row_entries(G[-2], y[-1], x, dtype)
    assert min(ent) == m
    jstar = max(j for j in range(lx) if ent[j] == m)
    assert all(Mo[-1, j] == 4 for j in range(jstar + 1, lx))
    assert Mo[-1, jstar] in ((3, 255) if jstar == 0 else (1, 3))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("w", [3, 4, 7, 16])
def test_open_end_swath_left_columns_closed_form(dtype, w):
    """|X| > W: the closed form with the diagonal free at j = kW, k >= 1;
    the early stop at swath starts leaves the distance as it is."""
    rng = np.random.default_rng(7 * w)
    hits_stop = 0
    hits_quirk = 0
    for _ in range(25):
        ly, lx = int(rng.integers(2, 9)), int(rng.integers(w + 1, 8 * w + 1))
        y, x = _pair(rng, dtype, ly, lx)
        G, _ = textbook_global(y, x, dtype)
        m = top_row_minimum(G[-2], y[-1], x, dtype, w)
        d, _, ab, _ = O.dtw_pair(y, x, os_=0, oe=1, w=w)
        assert d == f32_distance(m, ly, True, dtype), (ly, lx, d, m)
        hits_stop += ab >= 0
        hits_quirk += m < G[-1].min()
    assert hits_stop > 0
    if w <= 4:  # narrow swaths: the free diagonal lowers the minimum in some trials
        assert hits_quirk > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("w", [1024, 3, 4])
@pytest.mark.parametrize("seed", range(4))
def test_open_start_is_free_start_subsequence_dtw(dtype, w, seed):
    """Row 0 is free, so rows 1.. are subsequence DTW of Y[1:] with a free
    start; open start alone does not depend on the swath width."""
    rng = np.random.default_rng(200 + seed)
    ly, lx = int(rng.integers(2, 10)), int(rng.integers(1, 14))
    y, x = _pair(rng, dtype, ly, lx)
    S, Ms = subsequence_dtw(y[1:], x, dtype)
    d, Mo, _, _ = O.dtw_pair(y, x, os_=1, oe=0, w=w, moves=True)
    assert d == f32_distance(S[-1, -1], ly, True, dtype), (d, S[-1, -1])
    assert Mo[0, 0] == 5 and (Mo[0, 1:] == 4).all()
    # the free start row's cells come from row 0 by a diagonal (column 0: up)
    assert Mo[1, 0] == 3 and (Mo[1, 1:] == 1).all()
    assert np.array_equal(Mo[2:], Ms[1:])


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("seed", range(6))
def test_open_is_subsequence_dtw_with_free_ends(dtype, seed):
    rng = np.random.default_rng(300 + seed)
    ly, lx = int(rng.integers(2, 10)), int(rng.integers(1, 14))
    y, x = _pair(rng, dtype, ly, lx)
    S, _ = subsequence_dtw(y[1:], x, dtype)
    d, _, ab, _ = O.dtw_pair(y, x, os_=1, oe=1)
    assert ab == -1  # no early stop with open start (dtw.hpp:74)
    assert d == f32_distance(S[-1].min(), ly, False, dtype), (d, S[-1].min())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("w", [4, 8])
def test_open_swath_left_columns_closed_form(dtype, w):
    rng = np.random.default_rng(400 + w)
    for _ in range(20):
        ly, lx = int(rng.integers(3, 9)), int(rng.integers(w + 1, 8 * w + 1))
        y, x = _pair(rng, dtype, ly, lx)
        S, _ = subsequence_dtw(y[1:], x, dtype)
        m = top_row_minimum(S[-2], y[-1], x, dtype, w)
        d = O.dtw_pair(y, x, os_=1, oe=1, w=w)[0]
        assert d == f32_distance(m, ly, False, dtype)


def _dtw_abr(y, x, w, abrogate, moves):
    L = O.lib()
    L.orc_dtw_pair_abr.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint64, C.c_int,
                                   C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    d = np.zeros(1, dtype=y.dtype)
    M = np.zeros((len(y), len(x)), dtype=np.uint8)
    ab = C.c_int(0)
    rc = L.orc_dtw_pair_abr(O.dtype_code(y), O._p(y), len(y), O._p(x), len(x), w, 0, 1, abrogate, O._p(M),
                            O._p(d), C.byref(ab))
    assert rc == 0
    return d[0], M, ab.value


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32])
@pytest.mark.parametrize("w", [4, 8, 16])
def test_open_end_stop_leaves_distance_and_counts(dtype, w):
    """The stop fires only where the previous swath's top-right value is its
    column minimum, which no later cell can undercut: the distance with the
    check switched off is the same, and so is the traceback's count."""
    rng = np.random.default_rng(500 + w)
    fired = 0
    for _ in range(40):
        ly, lx = int(rng.integers(2, w)), int(rng.integers(2 * w, 8 * w + 1))
        if dtype == np.uint32:
            y, x = rng.integers(0, 40, size=ly).astype(dtype), rng.integers(0, 40, size=lx).astype(dtype)
            x[: lx // 3] = y[-1]  # make the top row reach its minimum early
        else:
            y, x = _pair(rng, dtype, ly, lx)
        d1, M1, ab1 = _dtw_abr(y, x, w, 1, True)
        d0, M0, ab0 = _dtw_abr(y, x, w, 0, True)
        assert ab0 == -1
        assert d1 == d0
        if ab1 >= 0:
            fired += 1
            if np.issubdtype(dtype, np.floating):
                r1 = O.traceback_full(M1, 0, y, len(x))
                r0 = O.traceback_full(M0, 0, y, len(x))
                assert r1[0] == 0 and r0[0] == 0
                assert np.array_equal(r1[2], r0[2]) and np.array_equal(r1[1], r0[1])
    assert fired >= 5
