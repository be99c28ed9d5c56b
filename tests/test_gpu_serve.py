"""Serving loop: every job's result is the same run alone, mixed with other
jobs on the device, and through the plain per-job path; admission respects
the HBM budget."""
import numpy as np
import pytest

from opendba_amd import native as N
from opendba_amd import serve as S
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _jobs(seed=3, n_jobs=6):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_jobs):
        n = int(rng.integers(4, 13))
        lens = rng.integers(300, 3000, size=n)
        if k % 2:
            seqs = [WL.squiggle(int(L), rng).astype(np.float32) for L in lens]
        else:
            seqs = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in lens]
        jobs.append(dict(arrival=0.0, seqs=seqs, open_end=int(k % 3 == 1)))
    return jobs


def _same(a, b):
    assert a["medoid"] == b["medoid"] and a["rounds"] == b["rounds"]
    assert np.array_equal(a["centroid"].view(np.uint32), b["centroid"].view(np.uint32))


def test_server_results_equal_alone_mixed_and_plain():
    jobs = _jobs()
    with N.Context(0) as ctx:
        plain = [ctx.perform_dba(j["seqs"], 0, j["open_end"], norm=1, max_rounds=40) for j in jobs]
    alone = []
    with S.Server(0, 0, 1) as srv:
        for k, j in enumerate(jobs):
            jid, dt, ml = srv.submit(j["seqs"], 0, j["open_end"], 1, 40)
            assert srv.wait(jid) == N.JOB_DONE
            alone.append(srv.result(jid, dt, ml))
            print(f"alone job {k}: rounds {alone[-1]['rounds']}", flush=True)
    with S.Server(0, 0, 4) as srv:
        hs = [srv.submit(j["seqs"], 0, j["open_end"], 1, 40) for j in jobs]
        mixed = []
        for jid, dt, ml in hs:
            assert srv.wait(jid) == N.JOB_DONE
            mixed.append(srv.result(jid, dt, ml))
    print("mixed done", flush=True)
    for p, a, m in zip(plain, alone, mixed):
        assert a["rc"] == 0 and m["rc"] == 0
        _same(a, m)
        _same(dict(medoid=p["medoid"], rounds=p["rounds"], centroid=p["centroid"]), a)
    # the mixed run overlapped: some job was admitted before another finished
    spans = sorted((m["t_admit"], m["t_done"]) for m in mixed)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))


def test_server_budget_serialises_admission():
    """A budget below every job's working set admits a job only when nothing
    else runs."""
    jobs = _jobs(seed=5, n_jobs=4)
    with S.Server(0, 1 << 20, 4) as srv:
        hs = [srv.submit(j["seqs"], 0, j["open_end"], 1, 10) for j in jobs]
        res = []
        for jid, dt, ml in hs:
            srv.wait(jid)
            res.append(srv.result(jid, dt, ml))
        assert srv.stats()["running"] == 0
    for a, b in zip(res, res[1:]):
        assert b["t_admit"] >= a["t_done"]


def test_trace_runner_plain_and_served_agree():
    jobs = WL.poisson_jobs(n_jobs=5, rate=50.0, seed=11, scale=0.05)
    a = S.run_trace(jobs, "plain", max_rounds=20)
    b = S.run_trace(jobs, "served", slots=4, max_rounds=20)
    for x, y in zip(a["results"], b["results"]):
        _same(x, y)
    assert b["jobs_per_s"] > 0 and a["p99_s"] >= a["p50_s"]
