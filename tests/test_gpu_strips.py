"""The workgroup-per-pair all-vs-all schedule (column strips 64*C wide swept by
several waves, edges handed over through LDS-signalled slots) against the
oracle, in every mode and strip width, and against the one-wave-per-pair
schedule for the open-end abrogation flags it reconstructs after the fact.
"""
import contextlib
import os

import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N

pytestmark = pytest.mark.gpu
MODES = [(0, 0), (0, 1), (1, 0), (1, 1)]


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@contextlib.contextmanager
def schedule(name):
    old = os.environ.get("ODBA_PAIR_SCHEDULE")
    os.environ["ODBA_PAIR_SCHEDULE"] = name
    try:
        yield
    finally:
        if old is None:
            del os.environ["ODBA_PAIR_SCHEDULE"]
        else:
            os.environ["ODBA_PAIR_SCHEDULE"] = old


def _walks(rng, lens, dtype=np.float32):
    return [np.cumsum(rng.normal(size=int(L))).astype(dtype) for L in lens]


def _run(ctx, seqs, os_, oe, sched):
    with schedule(sched):
        sset = ctx.seqset(seqs)
        try:
            got, flags, _ = sset.all_pairs(os_, oe, flags=True)
        finally:
            sset.close()
    return got, flags


# strip width follows the longest sequence: <= 1024 -> C=1, <= 4096 -> C=4, else C=16
CASES = {"c1": (200, 400, 12), "c4": (1100, 3000, 8), "c16": (4200, 9000, 6)}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("os_,oe", MODES)
def test_strip_schedule_bitexact(ctx, case, os_, oe):
    lo, hi, n = CASES[case]
    rng = np.random.default_rng(10 * list(CASES).index(case) + 2 * os_ + oe)
    lens = rng.integers(lo, hi + 1, size=n)
    seqs = O.znorm(_walks(rng, lens))[0]
    print(f"strips {case} os={os_} oe={oe} lens {lens.min()}..{lens.max()}", flush=True)
    got, flags = _run(ctx, seqs, os_, oe, "strip")
    exp = O.all_pairs(seqs, os_, oe)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    _, flags_w = _run(ctx, seqs, os_, oe, "wave")
    assert np.array_equal(flags, flags_w)


def test_strip_schedule_open_end_abrogation(ctx):
    rng = np.random.default_rng(7)
    seqs = O.znorm(_walks(rng, [150, 300, 900, 5000, 7000, 9000]))[0]
    print("strips abrogation", flush=True)
    got, flags = _run(ctx, seqs, 0, 1, "strip")
    exp = O.all_pairs(seqs, 0, 1)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    _, flags_w = _run(ctx, seqs, 0, 1, "wave")
    assert (flags_w & 1).sum() > 0
    assert np.array_equal(flags, flags_w)


def test_strip_schedule_uneven_lengths(ctx):
    """Pairs of one launch with 1 to 80 strips: waves without a strip in a
    pair wait at the workgroup barrier."""
    rng = np.random.default_rng(8)
    lens = [64, 65, 700, 1023, 1024, 1025, 2048, 3000, 5000]
    seqs = O.znorm(_walks(rng, lens))[0]
    for os_, oe in MODES:
        print(f"strips uneven os={os_} oe={oe}", flush=True)
        got, _ = _run(ctx, seqs, os_, oe, "strip")
        exp = O.all_pairs(seqs, os_, oe)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (os_, oe)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32])
def test_strip_schedule_integers(ctx, dtype):
    rng = np.random.default_rng(9)
    lo = -50 if dtype == np.int32 else 0
    seqs = [rng.integers(lo, 60, size=L).astype(dtype) for L in (300, 700, 1500, 2600)]
    for os_, oe in MODES:
        print(f"strips {np.dtype(dtype).name} os={os_} oe={oe}", flush=True)
        got, _ = _run(ctx, seqs, os_, oe, "strip")
        exp = O.all_pairs(seqs, os_, oe)
        assert np.array_equal(got, exp), (os_, oe)


def test_small_workload_both_schedules_match(ctx):
    """The 32-series small workload (496 pairs) on both schedules equals the
    oracle."""
    from opendba_amd import workloads as WL
    seqs = WL.random_walks(32, 200, 400, seed=0)
    seqs, _ = WL.stable_length_sort(O.znorm(seqs)[0])
    exp = O.all_pairs(seqs, 0, 0)
    for sched in ("wave", "strip"):
        got, _ = _run(ctx, seqs, 0, 0, sched)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), sched


@contextlib.contextmanager
def path_schedule(name):
    old = os.environ.get("ODBA_PATH_SCHEDULE")
    os.environ["ODBA_PATH_SCHEDULE"] = name
    try:
        yield
    finally:
        if old is None:
            del os.environ["ODBA_PATH_SCHEDULE"]
        else:
            os.environ["ODBA_PATH_SCHEDULE"] = old


@pytest.mark.parametrize("os_,oe", MODES)
@pytest.mark.parametrize("ly,lx", [(300, 2100), (2500, 4100), (1030, 1025), (5000, 3300)])
def test_path_strips_step_matrix(ctx, os_, oe, ly, lx):
    """Path jobs swept by a workgroup of waves (one 1024-column swath each)
    give the oracle's step matrix, open-end stop included."""
    rng = np.random.default_rng(ly + lx + 2 * os_ + oe)
    y, x = O.znorm(_walks(rng, [ly, lx]))[0]
    print(f"path strips os={os_} oe={oe} {ly}x{lx}", flush=True)
    with path_schedule("strip"):
        d, M = ctx.dtw_path(y, x, os_, oe)
    ed, EM, _, _ = O.dtw_pair(y, x, os_=os_, oe=oe, moves=True)
    assert np.array_equal(M, EM)
    assert d == ed


@pytest.mark.parametrize("os_,oe", MODES)
def test_path_strips_dba_round(ctx, os_, oe):
    rng = np.random.default_rng(70 + 2 * os_ + oe)
    seqs = O.znorm(_walks(rng, [2100, 3600, 1500, 4200, 2900]))[0]
    cen = seqs[1].copy()
    print(f"dba round path strips os={os_} oe={oe}", flush=True)
    with path_schedule("strip"):
        got, counts, delta = ctx.dba_round(cen, seqs, os_, oe)
    rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=os_, oe=oe)
    assert rc == 0
    assert np.array_equal(counts, ecounts)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert delta == edelta
