"""DBA rounds replayed as hipGraphs against the same rounds run eagerly:
every replay's centroid, counts and delta bit for bit, with eager work on the
same context (which may move its scratch and force a new capture) and
synchronisations between the replays, in every mode and in int32."""
import contextlib
import os

import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N

pytestmark = pytest.mark.gpu
os.environ["ODBA_DBA_GRAPH_AFTER"] = "1"  # replay from the second round (the default waits longer)


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@contextlib.contextmanager
def graphs(on):
    old = os.environ.get("ODBA_DBA_GRAPH")
    os.environ["ODBA_DBA_GRAPH"] = "1" if on else "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["ODBA_DBA_GRAPH"]
        else:
            os.environ["ODBA_DBA_GRAPH"] = old


def _steps(ctx, seqs, cen, os_, oe, on, k, interleave=None):
    out = []
    with graphs(on), N.DbaSession(ctx, seqs, os_, oe) as s:
        s.set_centroid(cen)
        for i in range(k):
            out.append(s.step())
            if interleave is not None:
                interleave(i)
    return out


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_graph_replays_equal_eager_rounds(ctx, os_, oe):
    rng = np.random.default_rng(80 + 2 * os_ + oe)
    seqs = O.znorm([np.cumsum(rng.normal(size=int(L))).astype(np.float32)
                    for L in rng.integers(300, 2500, size=12)])[0]
    cen = seqs[3].copy()
    other = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in (3000, 9000, 20000)]

    def eager_work(i):  # all-vs-all on the same context between replays; larger sets grow its scratch
        ctx.all_pairs(other[: 2 + (i % 2)], 0, 1)

    print(f"graph vs eager os={os_} oe={oe}", flush=True)
    eager = _steps(ctx, seqs, cen, os_, oe, False, 7)
    replay = _steps(ctx, seqs, cen, os_, oe, True, 7, eager_work)
    for k, (a, b) in enumerate(zip(eager, replay)):
        assert a[0].tobytes() == b[0].tobytes(), k
        assert np.array_equal(a[1], b[1]), k
        assert a[2] == b[2], k


def test_graph_replays_int32(ctx):
    rng = np.random.default_rng(91)
    seqs = [rng.integers(-40, 40, size=int(L)).astype(np.int32) for L in rng.integers(200, 1500, size=9)]
    cen = seqs[0].copy()
    eager = _steps(ctx, seqs, cen, 0, 0, False, 5)
    replay = _steps(ctx, seqs, cen, 0, 0, True, 5)
    for a, b in zip(eager, replay):
        assert a[0].tobytes() == b[0].tobytes() and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_perform_dba_with_graphs_equals_oracle(ctx):
    rng = np.random.default_rng(92)
    seqs = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(400, 1200, size=10)]
    with graphs(True):
        r = ctx.perform_dba(seqs, 0, 1, norm=1, max_rounds=30)
    e = O.perform_dba(seqs, 0, 1, norm=1, max_rounds=30)
    assert r["rounds"] == e["rounds"] and r["rounds"] >= 3
    assert r["centroid"].tobytes() == e["centroid"].tobytes()
