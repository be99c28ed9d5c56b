"""The row-segment traceback against the cell-by-cell traceback (selected
with ODBA_TRACEBACK=steps) and the oracle: DBA rounds with recorded paths in
every mode, with the centroid on either axis, open-end abrogation, long
RIGHT runs across swaths and integer values; centroids, counts, deltas and
every recorded cell must agree."""
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
def walk(name):
    old = os.environ.get("ODBA_TRACEBACK")
    if name is None:
        os.environ.pop("ODBA_TRACEBACK", None)
    else:
        os.environ["ODBA_TRACEBACK"] = name
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("ODBA_TRACEBACK", None)
        else:
            os.environ["ODBA_TRACEBACK"] = old


def _round(ctx, seqs, cen, os_, oe, name):
    with walk(name), N.DbaSession(ctx, seqs, os_, oe) as s:
        s.set_centroid(cen)
        s.record_paths(True)
        out = s.step()
        paths = [s.path(k) for k in range(len(seqs))]
    return out, paths


def _same(a, b):
    (ca, na, da), pa = a
    (cb, nb, db), pb = b
    assert ca.tobytes() == cb.tobytes()
    assert np.array_equal(na, nb)
    assert da == db
    for x, y in zip(pa, pb):
        assert all(np.array_equal(u, v) for u, v in zip(x, y))


@pytest.mark.parametrize("os_,oe", MODES)
def test_rows_equal_steps_mixed_lengths(ctx, os_, oe):
    """Sequences shorter and longer than the centroid (the centroid goes on
    Y for longer ones in open-end mode), across several swaths."""
    rng = np.random.default_rng(100 + 2 * os_ + oe)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32)
                    for L in (90, 700, 1023, 1025, 2300, 3100, 5200, 64)])[0]
    cen = seqs[3].copy()
    print(f"traceback rows vs steps os={os_} oe={oe}", flush=True)
    _same(_round(ctx, seqs, cen, os_, oe, None), _round(ctx, seqs, cen, os_, oe, "steps"))


@pytest.mark.parametrize("os_,oe", MODES)
def test_rows_match_oracle_round(ctx, os_, oe):
    rng = np.random.default_rng(110 + 2 * os_ + oe)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32)
                    for L in rng.integers(100, 2600, size=9)])[0]
    cen = seqs[0].copy()
    (c, n, d), _ = _round(ctx, seqs, cen, os_, oe, None)
    rc, ec, en, ed = O.dba_update(cen, seqs, os_=os_, oe=oe)
    assert rc == 0
    assert np.array_equal(n, en) and c.tobytes() == ec.tobytes() and d == ed


def test_rows_open_end_abrogation_and_long_runs(ctx):
    """A short centroid against long sequences in open-end mode: the
    reference abandons the search after a few swaths (the rest of the top row
    is OPEN_RIGHT) and the path runs far along the top row."""
    rng = np.random.default_rng(120)
    base = np.cumsum(rng.normal(size=400))
    seqs = [np.concatenate([base + rng.normal(scale=0.01, size=400),
                            np.cumsum(rng.normal(size=int(L))) + 50]).astype(np.float32)
            for L in (3000, 6000, 9000)]
    cen = base.astype(np.float32)
    for oe in (0, 1):
        _same(_round(ctx, seqs, cen, 0, oe, None), _round(ctx, seqs, cen, 0, oe, "steps"))


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float64])
def test_rows_equal_steps_other_types(ctx, dtype):
    rng = np.random.default_rng(130)
    if np.dtype(dtype).kind in "iu":
        lo = -20 if dtype == np.int32 else 0
        seqs = [rng.integers(lo, 30, size=int(L)).astype(dtype) for L in (300, 900, 1500, 2100)]
    else:
        seqs = [np.cumsum(rng.normal(size=int(L))).astype(dtype) for L in (300, 900, 1500, 2100)]
    cen = seqs[1].copy()
    for os_, oe in MODES:
        _same(_round(ctx, seqs, cen, os_, oe, None), _round(ctx, seqs, cen, os_, oe, "steps"))
