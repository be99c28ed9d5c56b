"""Pins the CPU oracle before it is trusted as the parity checker.

Checks: the reference's own golden vectors and KATs (tests/openDBA_test.cu,
tests/good_files/text/test1), textbook brute-force DTW on tiny inputs, closed
forms and invariants. Each check is chosen so that a plausible slip in the
restatement (dropped term, wrong tie-break, transposed operand, wrong swath
rule) fails at least one of them.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

G = os.path.join(os.path.dirname(__file__), "golden")


def textbook_dtw(y, x, dtype):
    """Symmetric DTW with squared-difference local cost, min over the three
    predecessors, diag preferred then up on ties (the White-Neely order of
    dtw.hpp:202-227). Returns cost matrix and step codes (reference codes)."""
    ly, lx = len(y), len(x)
    D = np.zeros((ly, lx), dtype=dtype)
    M = np.zeros((ly, lx), dtype=np.uint8)
    BIG = np.finfo(dtype).max if np.issubdtype(dtype, np.floating) else np.iinfo(dtype).max
    for i in range(ly):
        for j in range(lx):
            d = dtype(y[i] - x[j])
            dd = dtype(d * d)
            if i == 0 and j == 0:
                D[i, j] = dd
                M[i, j] = 255
                continue
            if i == 0:
                D[i, j] = dtype(D[i, j - 1] + dd)
                M[i, j] = 2
                continue
            up = dtype(D[i - 1, j] + dd)
            right = dtype(D[i, j - 1] + dd) if j > 0 else BIG
            diag = dtype(D[i - 1, j - 1] + dd) if j > 0 else BIG
            if diag > up:
                v, m = (right, 2) if up > right else (up, 3)
            else:
                v, m = (right, 2) if diag > right else (diag, 1)
            D[i, j] = v
            M[i, j] = m
    return D, M


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_global_matches_textbook(dtype, seed):
    rng = np.random.default_rng(seed)
    ly, lx = rng.integers(1, 9), rng.integers(1, 11)
    if np.issubdtype(dtype, np.floating):
        y = rng.normal(size=ly).astype(dtype)
        x = rng.normal(size=lx).astype(dtype)
    else:
        y = rng.integers(-20, 20, size=ly).astype(dtype)
        x = rng.integers(-20, 20, size=lx).astype(dtype)
    D, M = textbook_dtw(y, x, dtype)
    for w in (1024, 3, 4):  # global mode must not depend on the swath width
        d, Mo, ab, q3 = O.dtw_pair(y, x, w=w, moves=True)
        if ly >= 2 or (lx - 1) % w % 3 == 0 or lx <= w and (lx - 1) % 3 == 0:
            exp = np.float32(np.sqrt(np.float32(D[-1, -1])))
            if np.issubdtype(dtype, np.integer):
                exp = dtype(np.trunc(exp))
            assert d == dtype(exp), (w, d, exp)
        if ly >= 2:
            assert np.array_equal(Mo, M), w


def test_ties_prefer_diag_then_up():
    # constant sequences: every candidate ties -> DIAG everywhere off the edges
    y = np.ones(5, np.float32)
    x = np.ones(7, np.float32)
    d, M, _, _ = O.dtw_pair(y, x, moves=True)
    assert d == 0
    assert (M[1:, 1:] == 1).all() and (M[1:, 0] == 3).all() and (M[0, 1:] == 2).all() and M[0, 0] == 255


def test_transpose_symmetry_global():
    rng = np.random.default_rng(5)
    for _ in range(10):
        y = rng.normal(size=rng.integers(2, 40)).astype(np.float64)
        x = rng.normal(size=rng.integers(2, 40)).astype(np.float64)
        a = O.dtw_pair(y, x, w=7)[0]
        b = O.dtw_pair(x, y, w=5)[0]
        assert a == b


def test_open_end_le_global_and_self_zero():
    rng = np.random.default_rng(3)
    for _ in range(10):
        y = rng.normal(size=rng.integers(2, 30)).astype(np.float32)
        x = rng.normal(size=rng.integers(30, 60)).astype(np.float32)
        g = O.dtw_pair(y, x)[0]
        oe = O.dtw_pair(y, x, oe=1)[0] * np.float32(len(y))
        assert oe <= g * (1 + 1e-6)
        assert O.dtw_pair(y, y)[0] == 0


def test_open_end_swath_left_quirk_changes_result():
    """At swath-left columns the open-end top row takes the diagonal without
    the local cost (dtw.hpp:175-186): with a tiny swath width the distance
    differs from the W=1024 one on data built to expose it."""
    rng = np.random.default_rng(11)
    hits = 0
    for _ in range(40):
        y = rng.normal(size=6).astype(np.float64)
        x = rng.normal(size=30).astype(np.float64)
        if O.dtw_pair(y, x, oe=1, w=4)[0] != O.dtw_pair(y, x, oe=1, w=1024)[0]:
            hits += 1
    assert hits > 0


def test_traceback_counts_equal_path_minus_open():
    rng = np.random.default_rng(7)
    for oe in (0, 1):
        for os_ in (0, 1):
            y = rng.normal(size=23).astype(np.float64)
            x = rng.normal(size=41).astype(np.float64)
            _, M, _, _ = O.dtw_pair(y, x, os_=os_, oe=oe, moves=True)
            rc, sums, counts, pl = O.traceback_full(M, 0, y, len(x))
            assert rc == 0
            # walk the path in Python and count non-open steps
            i, j = len(y) - 1, len(x) - 1
            n = 0
            while M[i, j] not in (255, 5):
                if M[i, j] != 4:
                    n += 1
                di, dj = {1: (-1, -1), 2: (0, -1), 3: (-1, 0), 4: (0, -1)}[M[i, j]]
                i, j = i + di, j + dj
            if M[i, j] != 5:
                n += 1
            assert counts.sum() == n


def test_golden_test1_double_exact():
    s1 = O.read_text_series(os.path.join(G, "test1/seq1_test.txt"), np.float64)
    s2 = O.read_text_series(os.path.join(G, "test1/seq2_test.txt"), np.float64)
    exp = O.read_tokens(os.path.join(G, "test1/output_results.txt"))
    r = O.perform_dba([s1, s2], norm=0)
    assert r["medoid"] == 1
    assert np.array_equal(r["centroid"], exp)


def test_golden_test1_znorm_exact():
    z1 = O.read_tokens(os.path.join(G, "test1/seq1_test_znorm.txt"))
    z2 = O.read_tokens(os.path.join(G, "test1/seq2_test_znorm.txt"))
    exp = O.read_tokens(os.path.join(G, "test1/output_results_znorm.txt"))
    r = O.perform_dba([z1, z2], norm=0)
    assert np.array_equal(r["centroid"], exp)


def test_znorm_population_sd():
    s1 = O.read_text_series(os.path.join(G, "test1/seq1_test.txt"), np.float64)
    s2 = O.read_text_series(os.path.join(G, "test1/seq2_test.txt"), np.float64)
    (z, mu, sg) = O.znorm([s1, s2])
    for k, f in enumerate(["seq1_test_znorm.txt", "seq2_test_znorm.txt"]):
        ref = O.read_tokens(os.path.join(G, "test1", f))
        np.testing.assert_allclose(z[k], ref, rtol=0, atol=1e-14)
        np.testing.assert_allclose(mu[k], s1.mean() if k == 0 else s2.mean(), rtol=1e-15)
        np.testing.assert_allclose(sg[k], (s1 if k == 0 else s2).std(ddof=0), rtol=1e-14)


def test_znorm_constant_series_sigma_one():
    (z, mu, sg) = O.znorm([np.full(3000, 7.5, np.float32)])
    assert sg[0] == 0 and mu[0] == 7.5 and (z[0] == 0).all()


def test_simple_constant_centroid():
    s = [O.read_text_series(os.path.join(G, "test1", f"simple{k}_test.txt"), np.float64) for k in (1, 2)]
    exp = O.read_tokens(os.path.join(G, "test1/output_results_simple.txt"))
    r = O.perform_dba(s, norm=0)
    assert np.array_equal(r["centroid"], exp[: len(r["centroid"])])


@pytest.mark.parametrize("name", ["same", "different", "three", "truncated", "large"])
def test_reference_kats(name):
    K = json.load(open(os.path.join(G, "kat_expected.json")))[name]
    seqs = [O.read_text_series(os.path.join(G, "test2", f), np.float32) for f in K["files"]]
    r = O.perform_dba(seqs, oe=K["open_end"], norm=0)
    c = r["centroid"]
    assert len(c) == K["n"]
    for i, v in K["exact"].items():
        assert c[int(i)] == np.float32(v)
    for i, v in K["round3"].items():
        assert np.float32(float("%.3f" % c[int(i)])) == np.float32(v), (i, c[int(i)], v)


def test_stripe_equals_full_in_global_mode():
    rng = np.random.default_rng(21)
    for dtype in (np.float32, np.float64):
        C = rng.normal(size=37).astype(dtype)
        seqs = [rng.normal(size=n).astype(dtype) for n in (30, 45, 12)]
        a = O.dba_update(C, seqs, path_policy=1)
        b = O.dba_update(C, seqs, path_policy=2)
        assert a[0] == 0 and b[0] == 0
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_medoid_rules():
    # two members: the longer one, the second on a tie (dba.hpp:336-339)
    m, _ = O.medoid_single(np.array([1.0], np.float32), [5, 5])
    assert m == 1
    m, _ = O.medoid_single(np.array([1.0], np.float32), [6, 5])
    assert m == 0
    # three members: first strict minimum of sum of squares
    d = np.array([1.0, 2.0, 1.0], np.float32)  # (0,1)=1 (0,2)=2 (1,2)=1
    m, sos = O.medoid_single(d, [3, 3, 3])
    assert m == 1 and list(sos) == [5.0, 2.0, 5.0]


def test_kseg_parameters_match_reference_formulas():
    # ds = ceil(minseg/3), 256 threads, max_k = DIV_ROUNDUP(256, minseg/ds) in float (segmentation.hpp:347-353)
    assert O.kseg_params(8) == (3, 768, 96)
    assert O.kseg_params(4) == (2, 512, 128)
    assert O.kseg_params(3) == (1, 256, 86)


def test_kseg_constant_and_step_signals():
    """Closed forms: a constant signal is one level per chunk, merged across
    chunk borders; a clean two-level step yields exactly its two levels."""
    flat = [np.full(3000, 90.0, np.float32)]
    ev = O.kseg(flat, 8)[0]
    assert len(ev) >= 1 and np.all(ev == 90.0)
    step = np.concatenate([np.full(300, 80.0), np.full(300, 110.0)]).astype(np.float32)
    ev = O.kseg([step], 8)[0]
    assert set(np.unique(ev)) <= {80.0, 110.0, 95.0}
    assert ev[0] == 80.0 and ev[-1] == 110.0


def test_kseg_events_are_segment_medians_of_input_values():
    from opendba_amd import workloads as WL
    raw = WL.short_to_float(WL.squiggles(4, 3000, 6000, seed=9))
    ev = O.kseg(raw, 8)
    for r, e in zip(raw, ev):
        assert 0 < len(e) < len(r)
        # medians of integer-valued samples are integers or half-integers; the
        # chunk-border average can add quarter steps
        assert np.all(np.abs(e * 4 - np.round(e * 4)) == 0)
        assert e.min() >= r.min() and e.max() <= r.max()


# chopPrefixFromSequences (dba.hpp:1192-1395)

def test_prefix_chop_exact_prefix():
    """The sequence starts with the prefix itself, then far-away values: the
    path is the diagonal and the top row is open-right after it."""
    rng = np.random.default_rng(31)
    y = rng.permutation(50).astype(np.float32)          # distinct values
    tail = (100.0 + rng.permutation(300)).astype(np.float32)
    chop, hist = O.prefix_chop(y, np.concatenate([y, tail]))
    assert chop == len(y) - 1
    assert np.array_equal(hist, np.ones(len(y), dtype=np.int32))


def test_prefix_chop_doubled_values_hand_derived():
    """prefix [0, 10, 20], sequence [0, 0, 10, 10, 20, 20, 100, 100, 100]:
    walking back, (2,8)..(2,5) are OPEN_RIGHT, (2,4) is DIAG (chop = 4), then
    (1,3) RIGHT, (1,2) DIAG, (0,1) RIGHT, (0,0) NIL."""
    y = np.array([0, 10, 20], dtype=np.float32)
    x = np.array([0, 0, 10, 10, 20, 20, 100, 100, 100], dtype=np.float32)
    chop, hist = O.prefix_chop(y, x)
    assert chop == 4
    assert hist.tolist() == [2, 2, 1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_prefix_chop_histogram_is_the_dba_count(seed):
    """The histogram is the per-row count of the DBA traceback with the
    prefix on Y (every cell but the top row's open-right run, plus (0,0)),
    and the chop position is the first counted column of the top row."""
    rng = np.random.default_rng(seed)
    y = np.cumsum(rng.normal(size=40)).astype(np.float32)
    x = np.cumsum(rng.normal(size=int(rng.integers(60, 2600)))).astype(np.float32)
    chop, hist = O.prefix_chop(y, x)
    _, M, _, _ = O.dtw_pair(y, x, os_=0, oe=1, moves=True)
    rc, _, counts, _ = O.traceback_full(M, 1, x, len(y))
    assert rc == 0
    assert np.array_equal(hist, counts.astype(np.int32))
    top = M[len(y) - 1]
    assert all(top[j] == 4 for j in range(chop + 1, len(x))) and top[chop] != 4
