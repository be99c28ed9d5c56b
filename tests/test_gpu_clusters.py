"""Packed multi-cluster DBA (odba_dba_clusters): every cluster's centroid and
round count equal a per-cluster loop over the oracle's DBA round with the
reference's stop rule (dba.hpp:1084-1110), and the plain per-cluster
session path on the GPU."""
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


def _stop_loop(step, cen, max_rounds, open_mode):
    prev = np.zeros_like(cen)
    prev2 = np.zeros_like(cen)
    new, rounds = cen, 0
    for r in range(max_rounds):
        new, delta = step(cen)
        rounds = r + 1
        cen = new
        if delta == 0:
            break
        if open_mode:
            if r >= 1 and new.tobytes() == prev2.tobytes():
                break
            if r > 1:
                prev2 = prev.copy()
            if r > 0:
                prev = new.copy()
    return new, rounds


def _clusters(seed, sizes, lo=200, hi=900):
    rng = np.random.default_rng(seed)
    seqs, memb = [], []
    for k, m in enumerate(sizes):
        base = np.cumsum(rng.normal(size=hi))
        for _ in range(m):
            L = int(rng.integers(lo, hi))
            seqs.append((base[:L] + rng.normal(0, 0.3, size=L)).astype(np.float32))
            memb.append(k)
    seqs = O.znorm(seqs)[0]
    order = sorted(range(len(seqs)), key=lambda i: len(seqs[i]))
    seqs = [seqs[i] for i in order]
    memb = np.array([memb[i] for i in order], dtype=np.int32)
    medoids = []
    for k in range(len(sizes)):
        idx = np.nonzero(memb == k)[0]
        medoids.append(int(idx[len(idx) // 2]))
    return seqs, memb, np.array(medoids, dtype=np.int32)


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 1)])
def test_packed_clusters_equal_oracle_and_plain(ctx, os_, oe):
    seqs, memb, med = _clusters(31 + 2 * os_ + oe, [5, 1, 7, 3, 2])
    R = 15
    print(f"clusters os={os_} oe={oe}: {len(seqs)} sequences", flush=True)
    cents, rounds = ctx.dba_clusters(seqs, memb, med, os_, oe, max_rounds=R)
    for k, m in enumerate(med):
        members = [seqs[i] for i in np.nonzero(memb == k)[0]]
        if len(members) == 1:
            assert rounds[k] == 0 and np.array_equal(cents[k], seqs[m])
            continue

        def ostep(c):
            rc, new, _, delta = O.dba_update(c, members, os_=os_, oe=oe)
            assert rc == 0
            return new, delta

        exp, er = _stop_loop(ostep, seqs[m].copy(), R, bool(os_ or oe))
        assert rounds[k] == er, (k, rounds[k], er)
        assert np.array_equal(cents[k].view(np.uint32), exp.view(np.uint32)), k
        with N.DbaSession(ctx, members, os_, oe) as s:
            s.set_centroid(seqs[m])

            def gstep(c):
                new, _, delta = s.step()
                return new, delta

            got, gr = _stop_loop(gstep, seqs[m].copy(), R, bool(os_ or oe))
        assert gr == er and np.array_equal(got.view(np.uint32), exp.view(np.uint32)), k


def test_packed_clusters_rescale_by_medoid():
    seqs, memb, med = _clusters(40, [4, 3])
    mu = np.linspace(-1.0, 1.0, len(seqs))
    sg = np.linspace(0.5, 2.0, len(seqs))
    with N.Context(0) as c:
        a, _ = c.dba_clusters(seqs, memb, med, 0, 0, max_rounds=5)
        b, _ = c.dba_clusters(seqs, memb, med, 0, 0, max_rounds=5, mu=mu, sigma=sg)
    for k, m in enumerate(med):
        exp = (mu[m] + a[k].astype(np.float64) * sg[m]).astype(np.float32)
        assert np.array_equal(b[k], exp)
