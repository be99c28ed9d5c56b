"""All-vs-all autotuner: every candidate schedule (one wave per pair in both
packing orders; bands of every strip and band width) gives the oracle's
distances bit for bit, and the tuned choice is remembered for the set's
shape and used by later calls."""
import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu

CANDIDATES = [dict(kind=0, order=0), dict(kind=0, order=1)] + [
    dict(kind=1, band_c=c, band_waves=w, ring=r) for c, w, r in
    ((1, 2, 2048), (1, 8, 128), (4, 2, 256), (4, 4, 2048), (16, 2, 2048))]


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_every_schedule_matches_oracle(ctx, os_, oe):
    rng = np.random.default_rng(60 + 2 * os_ + oe)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32)
                    for L in rng.integers(200, 1400, size=14)])[0]
    exp = O.all_pairs(seqs, os_, oe)
    s = ctx.seqset(seqs)
    try:
        for cand in CANDIDATES:
            print(f"schedule {cand} os={os_} oe={oe}", flush=True)
            s.set_schedule(cand, os_, oe)
            got, _, _ = s.all_pairs(os_, oe)
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), cand
        s.set_schedule(None, os_, oe)
    finally:
        s.close()


def test_tuner_choice_is_remembered_and_exact(ctx):
    seqs = WL.random_walks(32, 200, 400, seed=0)
    seqs, _ = WL.stable_length_sort(O.znorm(seqs)[0])
    exp = O.all_pairs(seqs, 0, 0)
    s = ctx.seqset(seqs)
    try:
        best, ms, tried = s.tune(0, 0, reps=3)
        print(f"tuned {best} {ms:.4f} ms over {tried} candidates", flush=True)
        assert tried >= 3 and ms > 0
        assert s.schedule(0, 0) == best
        got, _, _ = s.all_pairs(0, 0)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    finally:
        s.close()
    s2 = ctx.seqset(seqs)  # same shape: the choice carries over
    try:
        assert s2.schedule(0, 0) == best
    finally:
        s2.close()


def test_integer_sets_that_may_wrap_refuse_bands(ctx):
    rng = np.random.default_rng(3)
    seqs = [rng.integers(-2**30, 2**30, size=300).astype(np.int32) for _ in range(4)]
    s = ctx.seqset(seqs)
    try:
        with pytest.raises(N.OdbaError):
            s.set_schedule(dict(kind=1, band_c=1, band_waves=2, ring=2048))
    finally:
        s.close()
