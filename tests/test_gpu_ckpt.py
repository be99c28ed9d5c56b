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
