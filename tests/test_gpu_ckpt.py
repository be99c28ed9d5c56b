"""Checkpointed global paths (stripe-mode float jobs in global mode): the
DBA round and the recorded paths equal the full-matrix ones bit for bit,
with small blocks (many band and column-block hand-offs) and with the
default block size, and the round equals the oracle's."""
import contextlib
import os

import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    for k, v in kv.items():
        os.environ[k] = str(v)
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _round(ctx, seqs, cen, policy, rec=True):
    with N.DbaSession(ctx, seqs, 0, 0, path_policy=policy) as s:
        s.set_centroid(cen)
        if rec:
            s.record_paths(True)
        out = s.step()
        paths = [s.path(k) for k in range(len(seqs))] if rec else None
    return out, paths


def _same(a, b):
    (ca, na, da), pa = a
    (cb, nb, db), pb = b
    assert ca.tobytes() == cb.tobytes()
    assert np.array_equal(na, nb)
    assert da == db
    if pa is not None:
        for x, y in zip(pa, pb):
            assert all(np.array_equal(u, v) for u, v in zip(x, y))


@pytest.mark.parametrize("rshift,ew", [(7, 2), (8, 2), (10, 16)])
def test_ckpt_equals_full_small_blocks(ctx, rshift, ew):
    rng = np.random.default_rng(140 + rshift)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32)
                    for L in (300, 1500, 2047, 2049, 4100, 6000, 9000)])[0]
    cen = seqs[4].copy()
    print(f"ckpt R=2^{rshift} EW={ew}", flush=True)
    full = _round(ctx, seqs, cen, N.PATH_FULL)
    with env(ODBA_CKPT_RSHIFT=rshift, ODBA_CKPT_EW=ew):
        ck = _round(ctx, seqs, cen, N.PATH_STRIPE)
    _same(full, ck)


def test_ckpt_default_blocks_long_sequences(ctx):
    """Sequences longer than one default block (16,384 rows, 16,384
    columns) on both axes."""
    rng = np.random.default_rng(150)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in (20000, 34000, 17000)])[0]
    cen = seqs[1].copy()
    print("ckpt default blocks", flush=True)
    full = _round(ctx, seqs, cen, N.PATH_FULL, rec=False)
    ck = _round(ctx, seqs, cen, N.PATH_STRIPE, rec=False)
    _same(full, ck)


def test_ckpt_round_matches_oracle(ctx):
    rng = np.random.default_rng(160)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(500, 3000, 6)])[0]
    cen = seqs[2].copy()
    with env(ODBA_CKPT_RSHIFT=7, ODBA_CKPT_EW=2):
        (c, n, d), _ = _round(ctx, seqs, cen, N.PATH_STRIPE, rec=False)
    rc, ec, en, ed = O.dba_update(cen, seqs, os_=0, oe=0, path_policy=2)
    assert rc == 0
    assert np.array_equal(n, en) and c.tobytes() == ec.tobytes() and d == ed


def _round_mode(ctx, seqs, cen, os_, oe, rec=True):
    with N.DbaSession(ctx, seqs, os_, oe, path_policy=N.PATH_STRIPE) as s:
        s.set_centroid(cen)
        if rec:
            s.record_paths(True)
        out = s.step()
        paths = [s.path(k) for k in range(len(seqs))] if rec else None
    return out, paths


OPEN_MODES = [(0, 1), (1, 0), (1, 1)]


@pytest.mark.parametrize("rshift", [6, 8])
@pytest.mark.parametrize("os_,oe", OPEN_MODES)
def test_open_stripe_ckpt_equals_stripe_recompute(ctx, os_, oe, rshift):
    """Open-mode stripe paths with checkpointed bands of 2^rshift rows: the
    round and every recorded path equal the recompute of each stripe from
    row 0 (ODBA_CKPT=0) and the round equals the oracle's stripe mode.
    Lengths put stripe top rows on band starts (64k + 1 rows: the top row
    alone in its band) and, in open end, sequences on both sides of the
    centroid's length (flipped pairs)."""
    rng = np.random.default_rng(170 + 4 * os_ + 2 * oe + rshift)
    lens = [193, 257, 641, 1100, 2049, 3300, 4200, 5000]
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in lens])[0]
    cen = seqs[5].copy()
    print(f"open stripe ckpt os={os_} oe={oe} R=2^{rshift}", flush=True)
    with env(ODBA_CKPT=0):
        ref = _round_mode(ctx, seqs, cen, os_, oe)
    with env(ODBA_CKPT_RSHIFT=rshift):
        ck = _round_mode(ctx, seqs, cen, os_, oe)
    _same(ref, ck)
    (c, n, d), _ = ck
    rc, ec, en, ed = O.dba_update(cen, seqs, os_=os_, oe=oe, path_policy=2)
    assert rc == 0
    assert np.array_equal(n, en) and c.tobytes() == ec.tobytes() and d == ed


@pytest.mark.parametrize("os_,oe", OPEN_MODES)
def test_open_stripe_ckpt_default_bands_many_stripes(ctx, os_, oe):
    """Default bands (16,384 rows) with sequences longer than one band and
    than sixteen stripes (one claim of the forward pass): equal to the
    recompute from row 0."""
    rng = np.random.default_rng(180 + 2 * os_ + oe)
    lens = [17000, 21000, 9000]
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in lens])[0]
    cen = seqs[0].copy()
    print(f"open stripe ckpt default bands os={os_} oe={oe}", flush=True)
    with env(ODBA_CKPT=0):
        ref = _round_mode(ctx, seqs, cen, os_, oe, rec=False)
    ck = _round_mode(ctx, seqs, cen, os_, oe, rec=False)
    _same(ref, ck)
