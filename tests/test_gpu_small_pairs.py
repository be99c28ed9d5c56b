"""The lock-step small-pair all-vs-all (float, lengths 64..1024, a workgroup
of waves per pair) against the oracle, bit for bit: lengths at every
strip and lane boundary of the three strip widths, both sub-block lengths,
the four modes, both packing orders; the small method's set under the
built-in choice; refusals of sets the kernel cannot take."""
import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu

BOUNDARY = [64, 65, 127, 128, 129, 200, 255, 256, 257, 383, 384, 385, 511, 512, 513, 640, 767, 768, 769, 900, 1023, 1024]
BOUNDARY_512 = [64, 65, 95, 96, 97, 127, 128, 129, 191, 192, 193, 255, 256, 257, 300, 319, 320, 321, 383, 384, 385,
                447, 448, 449, 500, 511, 512]


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _set(seed, lens):
    rng = np.random.default_rng(seed)
    return O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in lens])[0]


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_small_schedule_matches_oracle_at_boundaries(ctx, os_, oe):
    seqs = _set(70 + 2 * os_ + oe, BOUNDARY)
    exp = O.all_pairs(seqs, os_, oe)
    s = ctx.seqset(seqs)
    try:
        for C in (2, 4):
            waves = -(-max(BOUNDARY) // (64 * C))
            for order, ring in ((0, 16), (1, 32)):
                cand = dict(kind=2, order=order, band_c=C, band_waves=waves, ring=ring)
                print(f"small schedule {cand} os={os_} oe={oe}", flush=True)
                s.set_schedule(cand, os_, oe)
                got, fl, _ = s.all_pairs(os_, oe, flags=True)
                assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), cand
                assert not fl.any()
        s.set_schedule(None, os_, oe)
    finally:
        s.close()


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_small_schedule_one_column_per_lane(ctx, os_, oe):
    """C = 1: strips of 64 columns, up to 8 waves (lengths to 512)."""
    seqs = _set(90 + 2 * os_ + oe, BOUNDARY_512)
    exp = O.all_pairs(seqs, os_, oe)
    s = ctx.seqset(seqs)
    try:
        for order, ring in ((0, 16), (1, 32), (0, 32)):
            cand = dict(kind=2, order=order, band_c=1, band_waves=8, ring=ring)
            print(f"small schedule {cand} os={os_} oe={oe}", flush=True)
            s.set_schedule(cand, os_, oe)
            got, _, _ = s.all_pairs(os_, oe)
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), cand
    finally:
        s.close()


@pytest.mark.parametrize("lmax", [200, 400, 700])
def test_small_schedule_few_waves_per_pair(ctx, lmax):
    """Sets whose longest sequence needs 1, 2, 3 or 4 strips of 128 columns."""
    seqs = _set(lmax, np.random.default_rng(lmax).integers(64, lmax + 1, size=20))
    for os_, oe in ((0, 0), (0, 1)):
        exp = O.all_pairs(seqs, os_, oe)
        s = ctx.seqset(seqs)
        try:
            s.set_schedule(dict(kind=2, order=0, band_c=2, band_waves=-(-max(len(x) for x in seqs) // 128)), os_, oe)
            got, _, _ = s.all_pairs(os_, oe)
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
        finally:
            s.close()


def test_small_method_set_takes_the_small_schedule(ctx):
    seqs = WL.random_walks(32, 200, 400, seed=0)
    seqs, _ = WL.stable_length_sort(O.znorm(seqs)[0])
    s = ctx.seqset(seqs)
    try:
        sched = s.schedule(0, 0)
        assert sched["kind"] == 2 and sched["band_c"] == 4 and sched["ring"] == 32 and sched["order"] == 1
        got, _, _ = s.all_pairs(0, 0)
        assert np.array_equal(got.view(np.uint32), O.all_pairs(seqs, 0, 0).view(np.uint32))
    finally:
        s.close()
    assert np.array_equal(ctx.all_pairs(seqs, 0, 0).view(np.uint32), O.all_pairs(seqs, 0, 0).view(np.uint32))


def test_small_schedule_refusals(ctx):
    for seqs, cand in (
        (_set(1, [100, 1100]), dict(kind=2, order=0, band_c=2, band_waves=9)),      # longer than one swath
        (_set(2, [63, 300]), dict(kind=2, order=0, band_c=2, band_waves=3)),        # shorter than 64
        ([x.astype(np.float64) for x in _set(3, [100, 300])], dict(kind=2, order=0, band_c=2, band_waves=3)),
        (_set(4, [100, 300]), dict(kind=2, order=0, band_c=2, band_waves=4)),       # waves not ceil(300/128)
        (_set(5, [100, 300]), dict(kind=2, order=0, band_c=8, band_waves=1)),       # C not 1, 2 or 4
        (_set(6, [100, 600]), dict(kind=2, order=0, band_c=1, band_waves=10)),      # more than 8 waves
        (_set(7, [100, 300]), dict(kind=2, order=0, band_c=2, band_waves=3, ring=8)),  # sub-block not 16 or 32
    ):
        s = ctx.seqset(seqs)
        try:
            with pytest.raises(N.OdbaError):
                s.set_schedule(cand)
        finally:
            s.close()


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 1)])
def test_small_schedule_infinite_values(ctx, os_, oe):
    """Sequences holding +inf and -inf: without open start a strip whose
    columns hold +inf keeps the masked prologue (its steps before a lane's
    first row would compute inf - inf), one holding -inf does not. NaN
    samples are outside the parity claim of every float kernel (min-first
    drops NaN where the reference's comparisons keep it), so none is used."""
    seqs = _set(33 + 2 * os_ + oe, [70, 130, 200, 257, 300, 390, 400, 512])
    seqs[1][5] = np.inf
    seqs[3][200] = np.inf
    seqs[5][17] = -np.inf
    exp = O.all_pairs(seqs, os_, oe)
    s = ctx.seqset(seqs)
    try:
        for C, B in ((1, 16), (2, 32), (4, 16), (4, 32)):
            cand = dict(kind=2, order=1, band_c=C, band_waves=-(-512 // (64 * C)), ring=B)
            print(f"small schedule {cand} os={os_} oe={oe}", flush=True)
            s.set_schedule(cand, os_, oe)
            got, _, _ = s.all_pairs(os_, oe)
            nan = np.isnan(exp)
            assert np.array_equal(np.isnan(got), nan), cand
            assert np.array_equal(got[~nan].view(np.uint32), exp[~nan].view(np.uint32)), cand
    finally:
        s.close()
