"""The banded all-vs-all schedule (column strips 64*C wide, bands of strips
claimed by workgroups, edges handed over through rings inside a workgroup and
whole-column global slots between bands) against the oracle, in every mode,
strip width, band width and ring size, and against the one-wave-per-pair
schedule for the open-end abrogation flags it reconstructs after the fact.
"""
import contextlib
import os

import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N

pytestmark = pytest.mark.gpu
# a broken hand-off would end each wait after 10 s and report an error
os.environ.setdefault("ODBA_BAND_WAIT_S", "10")
MODES = [(0, 0), (0, 1), (1, 0), (1, 1)]


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    for k, v in kv.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = str(v)
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def schedule(name):
    return env(ODBA_PAIR_SCHEDULE=name)


# (strips per band, ring rows): the default (up to 16 strips per band, 2048
# rows), and narrow bands with small rings so that bands hand over between
# workgroups and the rings wrap many times.
BANDS = {"default": (None, None), "w2r128": (2, 128), "w3r256": (3, 256)}


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


@pytest.mark.parametrize("band", list(BANDS))
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("os_,oe", MODES)
def test_strip_schedule_bitexact(ctx, case, os_, oe, band):
    lo, hi, n = CASES[case]
    rng = np.random.default_rng(10 * list(CASES).index(case) + 2 * os_ + oe)
    lens = rng.integers(lo, hi + 1, size=n)
    seqs = O.znorm(_walks(rng, lens))[0]
    print(f"strips {case} os={os_} oe={oe} band {band} lens {lens.min()}..{lens.max()}", flush=True)
    w, r = BANDS[band]
    with env(ODBA_BAND_WAVES=w, ODBA_BAND_RING=r):
        got, flags = _run(ctx, seqs, os_, oe, "strip")
    exp = O.all_pairs(seqs, os_, oe)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    _, flags_w = _run(ctx, seqs, os_, oe, "wave")
    assert np.array_equal(flags, flags_w)


@pytest.mark.parametrize("band", list(BANDS))
def test_strip_schedule_open_end_abrogation(ctx, band):
    rng = np.random.default_rng(7)
    seqs = O.znorm(_walks(rng, [150, 300, 900, 5000, 7000, 9000]))[0]
    print(f"strips abrogation band {band}", flush=True)
    w, r = BANDS[band]
    with env(ODBA_BAND_WAVES=w, ODBA_BAND_RING=r):
        got, flags = _run(ctx, seqs, 0, 1, "strip")
    exp = O.all_pairs(seqs, 0, 1)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    _, flags_w = _run(ctx, seqs, 0, 1, "wave")
    assert (flags_w & 1).sum() > 0
    assert np.array_equal(flags, flags_w)


def test_band_schedule_one_strip_per_band(ctx):
    """Every strip its own band: each edge column between strips crosses
    workgroups through a global slot, and slots are reused many times over
    (far more claims than workgroups)."""
    rng = np.random.default_rng(12)
    seqs = O.znorm(_walks(rng, [5000, 6100, 7000, 8300]))[0]
    for os_, oe in MODES:
        print(f"bands of one strip os={os_} oe={oe}", flush=True)
        with env(ODBA_BAND_WAVES=1, ODBA_BAND_RING=128):
            got, _ = _run(ctx, seqs, os_, oe, "strip")
        exp = O.all_pairs(seqs, os_, oe)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (os_, oe)


def test_strip_schedule_uneven_lengths(ctx):
    """Pairs of one launch with 1 to 80 strips: waves without a strip in a
    pair wait at the workgroup barrier."""
    rng = np.random.default_rng(8)
    lens = [64, 65, 700, 1023, 1024, 1025, 2048, 3000, 5000]
    seqs = O.znorm(_walks(rng, lens))[0]
    for os_, oe in MODES:
        for band in ("default", "w2r128"):
            print(f"strips uneven os={os_} oe={oe} band {band}", flush=True)
            w, r = BANDS[band]
            with env(ODBA_BAND_WAVES=w, ODBA_BAND_RING=r):
                got, _ = _run(ctx, seqs, os_, oe, "strip")
            exp = O.all_pairs(seqs, os_, oe)
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (os_, oe, band)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32])
def test_strip_schedule_integers(ctx, dtype):
    rng = np.random.default_rng(9)
    lo = -50 if dtype == np.int32 else 0
    seqs = [rng.integers(lo, 60, size=L).astype(dtype) for L in (300, 700, 1500, 2600)]
    for os_, oe in MODES:
        print(f"strips {np.dtype(dtype).name} os={os_} oe={oe}", flush=True)
        with env(ODBA_BAND_WAVES=2, ODBA_BAND_RING=128):
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


# The band schedule of path jobs made an illegal memory access in its first
# open-end run (300 x 2100, global edge slots between workgroups); the cause is
# not found, the library refuses the schedule, and its cases stay skipped.
BAND_OFF = pytest.mark.skip(reason="path band schedule switched off: an open-end run faulted, cause not found")
# "swathband": bands of four whole swaths per workgroup; "swathband1": one
# swath per band, so every swath hands over to the next through a global slot.
PATH_SCHEDS = ["strip", "quarter", pytest.param("band", marks=BAND_OFF), "swathband", "swathband1"]


def path_schedule(name):
    if name == "swathband1":
        return env(ODBA_PATH_SCHEDULE="swathband", ODBA_PATH_BAND_WAVES=1)
    return env(ODBA_PATH_SCHEDULE=name, ODBA_PATH_BAND_WAVES=None)


@pytest.mark.parametrize("os_,oe", MODES)
@pytest.mark.parametrize("ly,lx", [(300, 2100), (2500, 4100), (1030, 1025), (5000, 3300), (90, 200), (1500, 400),
                                   (300, 700), (2000, 1000)])
@pytest.mark.parametrize("sched", PATH_SCHEDS)
def test_path_strips_step_matrix(ctx, os_, oe, ly, lx, sched):
    """Path jobs swept by a workgroup of waves (one 1024-column swath each,
    or one quarter swath each, whose four waves share the step words), or
    as bands (one swath per workgroup of four quarter-swath waves, swaths
    handed over between workgroups), give the oracle's step matrix, open-end
    stop included. The widths cover last swaths of 1 (unsplit), 4, 8, 12 and
    16 columns per lane."""
    rng = np.random.default_rng(ly + lx + 2 * os_ + oe)
    y, x = O.znorm(_walks(rng, [ly, lx]))[0]
    print(f"path {sched}s os={os_} oe={oe} {ly}x{lx}", flush=True)
    with path_schedule(sched):
        d, M = ctx.dtw_path(y, x, os_, oe)
    ed, EM, _, _ = O.dtw_pair(y, x, os_=os_, oe=oe, moves=True)
    assert np.array_equal(M, EM)
    assert d == ed


@pytest.mark.parametrize("os_,oe", MODES)
@pytest.mark.parametrize("sched", PATH_SCHEDS)
def test_path_strips_dba_round(ctx, os_, oe, sched):
    rng = np.random.default_rng(70 + 2 * os_ + oe)
    seqs = O.znorm(_walks(rng, [2100, 3600, 1500, 4200, 2900]))[0]
    cen = seqs[1].copy()
    print(f"dba round path {sched}s os={os_} oe={oe}", flush=True)
    with path_schedule(sched):
        got, counts, delta = ctx.dba_round(cen, seqs, os_, oe)
    rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=os_, oe=oe)
    assert rc == 0
    assert np.array_equal(counts, ecounts)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert delta == edelta


def test_band_schedule_long_pairs_default_bands(ctx):
    """Sequences long enough for the automatic choice (bands of 16 strips of
    1024 columns, several bands per pair, 2048-row rings) against the oracle,
    and the global distance's transposition symmetry on the reversed order."""
    from opendba_amd import workloads as WL
    seqs = WL.warped_motif_set(5, 30_000, 60_000, seed=3)
    print("long pairs, default bands", flush=True)
    got, _ = _run(ctx, seqs, 0, 0, "auto")
    exp = O.all_pairs(seqs, 0, 0)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    rev, _ = _run(ctx, seqs[::-1], 0, 0, "auto")
    n = len(seqs)
    iu = np.triu_indices(n, 1)
    R = np.zeros((n, n), dtype=np.float32)
    R[iu] = rev
    assert np.array_equal(R[::-1, ::-1].T[iu].view(np.uint32), got.view(np.uint32))


# Short sequences against long ones in open end: the reference stops a few
# swaths past the short sequence's length (dtw.hpp:74-129), and the strip
# schedules cancel every strip at or past that stop once it is decided. The
# step matrix returned by dtw_path starts from step words filled with an
# invalid code, so a read of a word the cancelled strips left unwritten would
# show as a wrong move.
# (700, 12600) never stops: every strip is swept, as before.
STOP_CASES = [(300, 9000), (64, 6000), (500, 12500), (2000, 16400), (1100, 7300), (150, 20000), (700, 12600)]


@pytest.mark.parametrize("ly,lx", STOP_CASES)
@pytest.mark.parametrize("sched", ["strip", "quarter", "swathband", "swathband1"])
def test_path_strips_open_end_stop_cancels(ctx, ly, lx, sched):
    rng = np.random.default_rng(ly * 7 + lx)
    y, x = O.znorm(_walks(rng, [ly, lx]))[0]
    ed, EM, eab, _ = O.dtw_pair(y, x, os_=0, oe=1, moves=True)
    K = (lx + 1023) // 1024
    assert eab == -1 or 1 <= eab < K, (eab, K)
    print(f"open-end stop {sched}s {ly}x{lx} (stop swath {eab} of {K})", flush=True)
    with path_schedule(sched):
        d, M = ctx.dtw_path(y, x, 0, 1)
    assert np.array_equal(M, EM)
    assert d == ed


@pytest.mark.parametrize("sched", ["strip", "quarter", "swathband", "swathband1"])
def test_path_strips_dba_round_long_centroid_open_end(ctx, sched):
    """A DBA round of short sequences against a long centroid (open end, no
    flip: Y = sequence), plus two sequences longer than the centroid (flipped:
    Y = centroid), against the oracle, twice in a row on the same context."""
    rng = np.random.default_rng(91)
    seqs = O.znorm(_walks(rng, [600, 900, 1500, 2500, 700, 14000, 15500]))[0]
    cen = O.znorm(_walks(rng, [13000]))[0][0]
    rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=0, oe=1)
    assert rc == 0
    print(f"dba round long centroid open end, {sched}s", flush=True)
    for _ in range(2):
        with path_schedule(sched):
            got, counts, delta = ctx.dba_round(cen, seqs, 0, 1)
        assert np.array_equal(counts, ecounts)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
        assert delta == edelta


@pytest.mark.parametrize("oe", [0, 1])
def test_path_swath_bands_chosen_for_a_long_centroid(ctx, oe):
    """Few jobs whose longest has more than 16 swaths take the swath-band
    schedule by default: a DBA round of short sequences against an
    18-swath centroid, plus one longer sequence (flipped in open end),
    against the oracle over three rounds."""
    rng = np.random.default_rng(93 + oe)
    seqs = O.znorm(_walks(rng, [700, 3000, 1800, 19000]))[0]
    cen = O.znorm(_walks(rng, [17500]))[0][0]
    print(f"swath bands by default oe={oe}", flush=True)
    with env(ODBA_PATH_SCHEDULE=None, ODBA_PATH_BAND_WAVES=None):
        for _ in range(3):
            got, counts, delta = ctx.dba_round(cen, seqs, 0, oe)
            rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=0, oe=oe)
            assert rc == 0
            assert np.array_equal(counts, ecounts)
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
            assert delta == edelta
            cen = exp


def _mixed_set(rng, dtype=np.float32):
    # 40 short series and three long ones: the largest pairs hold far more
    # than a resident wave's even share of the cells
    lens = list(rng.integers(300, 600, size=40)) + [6100, 8000, 9000]
    rng.shuffle(lens)
    if dtype == np.float32:
        return O.znorm(_walks(rng, lens))[0]
    lo = -50 if dtype == np.int32 else 0
    return [rng.integers(lo, 60, size=int(L)).astype(dtype) for L in lens]


@pytest.mark.parametrize("sched", ["mixed", "auto"])
@pytest.mark.parametrize("os_,oe", MODES)
def test_mixed_schedule_bitexact(ctx, os_, oe, sched):
    """The largest pairs as bands across the device, the rest one wave per
    pair: the oracle's distances, and the one-wave schedule's open-end stop
    flags. "auto" takes the mixed schedule by the default rule."""
    rng = np.random.default_rng(40 + 2 * os_ + oe)
    seqs = _mixed_set(rng)
    print(f"mixed schedule ({sched}) os={os_} oe={oe}", flush=True)
    got, flags = _run(ctx, seqs, os_, oe, sched)
    exp = O.all_pairs(seqs, os_, oe)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    _, flags_w = _run(ctx, seqs, os_, oe, "wave")
    assert np.array_equal(flags, flags_w)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32])
def test_mixed_schedule_integers(ctx, dtype):
    rng = np.random.default_rng(44)
    seqs = _mixed_set(rng, dtype)
    for os_, oe in MODES:
        print(f"mixed schedule {np.dtype(dtype).name} os={os_} oe={oe}", flush=True)
        got, _ = _run(ctx, seqs, os_, oe, "mixed")
        exp = O.all_pairs(seqs, os_, oe)
        assert np.array_equal(got, exp), (os_, oe)


@pytest.mark.parametrize("oe", [0, 1])
def test_path_swath_bands_more_than_1024_jobs(ctx, oe):
    """The swath-band schedule's claim table takes any number of jobs (the
    packed rounds of a server hand it thousands): a DBA round of 1,100 short
    sequences against a two-swath centroid, forced onto swath bands, against
    the oracle."""
    rng = np.random.default_rng(95 + oe)
    lens = rng.integers(80, 200, size=1100)
    seqs = O.znorm(_walks(rng, [int(L) for L in lens]))[0]
    cen = O.znorm(_walks(rng, [1500]))[0][0]
    print(f"swath bands over 1100 jobs oe={oe}", flush=True)
    with path_schedule("swathband"):
        got, counts, delta = ctx.dba_round(cen, seqs, 0, oe)
    rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=0, oe=oe)
    assert rc == 0
    assert np.array_equal(counts, ecounts)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert delta == edelta
