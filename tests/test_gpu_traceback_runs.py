"""DBA rounds whose paths hold long runs of one move, against the oracle.

The traceback walks a run of equal moves per step and skips the open-end
top row's OPEN_RIGHT stretch and row 0 in one go, writing one run record per
centroid position it passes. These inputs make those stretches hundreds to
thousands of cells long (more than one 64-lane store, across several
1024-column swaths): a centroid that matches the start of every sequence
followed by a long unrelated tail (open end: a long top-row run, with the
centroid on the rows or on the columns), a centroid that matches the end of
every sequence (open start: a long row-0 run), and sequences far longer
than the centroid (global: long UP runs and a long column-0 run).
Centroid, counts and delta must equal the oracle's bit for bit.
"""
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


def _walk(rng, n):
    return np.cumsum(rng.normal(size=n))


def _check(ctx, cen, seqs, os_, oe):
    rc, exp, ecounts, edelta = O.dba_update(cen, seqs, os_=os_, oe=oe)
    if rc:  # an integer centroid position with no count (0/0, which the reference traps)
        with pytest.raises(N.OdbaError) as ei:
            ctx.dba_round(cen, seqs, os_, oe)
        assert ei.value.code == rc
        return
    got, counts, delta = ctx.dba_round(cen, seqs, os_, oe)
    assert np.array_equal(counts, ecounts)
    if got.dtype.kind == "f":
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))
    else:
        assert np.array_equal(got, exp)
    assert delta == edelta


@pytest.mark.parametrize("os_,oe", MODES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_shared_start_long_tails(ctx, os_, oe, dtype):
    """A shared start followed by unrelated tails. (a) Sequences of the start
    plus 60-4,500 tail samples against a centroid of the start alone (open
    end: the centroid goes on the rows and the path runs 270-2,500 cells
    along its top row). (b) Sequences of the start alone against a centroid
    with a 2,500-sample tail (the centroid stays on the columns: the path
    runs along the sequences' last row through the tail)."""
    rng = np.random.default_rng(100 + 2 * os_ + oe)
    head = _walk(rng, 700)
    seqs = []
    for tail in (300, 900, 2600, 4500, 60):
        s = np.concatenate([head + rng.normal(0, 0.2, 700), head[-1] + _walk(rng, tail) * 3.0])
        seqs.append(s.astype(dtype))
    cen = (head + rng.normal(0, 0.2, 700)).astype(dtype)
    seqs = O.znorm(seqs)[0]
    cen = O.znorm([cen])[0][0]
    print(f"shared start (a) os={os_} oe={oe} {np.dtype(dtype).name}", flush=True)
    _check(ctx, cen, seqs, os_, oe)
    seqs = O.znorm([(head[:L] + rng.normal(0, 0.2, L)).astype(dtype) for L in (400, 550, 700)])[0]
    cen = O.znorm([np.concatenate([head, head[-1] + _walk(rng, 2500) * 3.0]).astype(dtype)])[0][0]
    print(f"shared start (b) os={os_} oe={oe} {np.dtype(dtype).name}", flush=True)
    _check(ctx, cen, seqs, os_, oe)


@pytest.mark.parametrize("os_,oe", MODES)
def test_shared_end_long_heads(ctx, os_, oe):
    """A centroid of an unrelated 2,000-sample head followed by a shared end,
    against sequences of the shared end alone (open start: the path meets
    row 0 about 2,000 columns from the corner)."""
    rng = np.random.default_rng(200 + 2 * os_ + oe)
    tail = _walk(rng, 600)
    h = _walk(rng, 2000) * 3.0
    cen = np.concatenate([h, h[-1] - tail[0] + tail]).astype(np.float32)
    seqs = [(tail[600 - L:] + rng.normal(0, 0.2, L)).astype(np.float32) for L in (300, 450, 600)]
    seqs = O.znorm(seqs)[0]
    cen = O.znorm([cen])[0][0]
    print(f"shared end os={os_} oe={oe}", flush=True)
    _check(ctx, cen, seqs, os_, oe)


@pytest.mark.parametrize("os_,oe", MODES)
@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_long_sequences_short_centroid(ctx, os_, oe, dtype):
    """Sequences 10-40 times longer than the centroid: long vertical runs
    (global) or a long top-row stretch over several swaths (open end, the
    centroid on the rows)."""
    rng = np.random.default_rng(300 + 2 * os_ + oe)
    if dtype == np.int32:
        seqs = [rng.integers(0, 40, size=L).astype(np.int32) for L in (2400, 5200, 9100)]
        cen = rng.integers(0, 40, size=230).astype(np.int32)
    else:
        seqs = O.znorm([_walk(rng, L).astype(np.float32) for L in (2400, 5200, 9100)])[0]
        cen = O.znorm([_walk(rng, 230).astype(np.float32)])[0][0]
    print(f"long sequences os={os_} oe={oe} {np.dtype(dtype).name}", flush=True)
    _check(ctx, cen, seqs, os_, oe)
