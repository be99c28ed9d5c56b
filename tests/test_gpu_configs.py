"""The benchmark configurations end to end at reduced sizes, each stage on the
device against the oracle (bit patterns compared):

- batched uneven: int16 squiggles (levels N(90,12), geometric dwell mean 8,
  noise sd 2) as float, z-norm, open-end all-vs-all, medoid, DBA to
  convergence (performDBA, dba.hpp:863-1189);
- segment-then-cluster: reads of shared molecules, 4k-40k samples,
  k-segmentation at minimum segment length 8 (segmentation.hpp:339-583),
  z-norm, open all-vs-all on the event series, complete-linkage clusters
  (host, pinned against SciPy in tests/test_cluster_cpu.py) and every
  cluster's consensus in packed rounds (dba.hpp:990-1110);
- serving: a seeded Poisson trace of jobs of mixed value kinds and modes,
  each job's medoid, round count and centroid from the serving loop equal
  to the oracle's performDBA of that job alone.
"""
import numpy as np
import pytest

import dropin_expect as E
import oracle_lib as O
from opendba_amd import native as N
from opendba_amd import serve as S
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", [21, 22])
def test_batched_uneven_dba_to_convergence(ctx, seed):
    seqs = WL.short_to_float(WL.squiggles(16, 800, 2500, seed=seed))
    print(f"batched uneven: 16 squiggles, seed {seed}", flush=True)
    got = ctx.perform_dba(seqs, 0, 1, norm=1, max_rounds=250)
    exp = O.perform_dba(seqs, 0, 1, norm=1, max_rounds=250)
    assert got["medoid"] == exp["medoid"] and got["rounds"] == exp["rounds"]
    assert np.array_equal(bits(got["centroid"]), bits(exp["centroid"]))


def test_segment_then_cluster_every_stage(ctx):
    raw16, _ = WL.squiggle_reads(5, 4, 4000, 40000, seed=5)
    raw = WL.short_to_float(raw16)
    print("segment-then-cluster: 20 reads of 5 molecules", flush=True)
    # k-segmentation to event medians
    ev = ctx.kseg(raw, 8)
    ev_exp = O.kseg(raw, 8, 0, 65536)
    assert all(np.array_equal(bits(a), bits(b)) for a, b in zip(ev, ev_exp))
    # z-norm of the event series
    z, mu, sg = ctx.znorm(ev)
    z_exp, mu_exp, sg_exp = O.znorm(ev_exp)
    assert all(np.array_equal(bits(a), bits(b)) for a, b in zip(z, z_exp))
    assert np.array_equal(mu, mu_exp) and np.array_equal(sg, sg_exp)
    # open all-vs-all on the length-sorted events
    order = E.stable_order([len(e) for e in z])
    z = [z[i] for i in order]
    lens = np.array([len(e) for e in z], dtype=np.uint64)
    sset = ctx.seqset(z)
    d, _, _ = sset.all_pairs(1, 1)
    sset.close()
    assert np.array_equal(bits(d), bits(O.all_pairs(z, 1, 1)))
    # clusters (K = 5 clusters of two or more members) and packed DBA
    memb, med = N.cluster(d, lens, 5.0)
    cents, rounds = ctx.dba_clusters(z, memb, med, 1, 1, max_rounds=250)
    multi = 0
    for c in range(len(med)):
        members = [z[i] for i in range(len(z)) if memb[i] == c]
        if len(members) < 2:
            continue
        multi += 1
        cen, r = E.consensus(members, z[med[c]], 1, 1)
        assert rounds[c] == r, c
        assert np.array_equal(bits(cents[c]), bits(cen)), c
    assert multi >= 3


def test_serving_trace_jobs_equal_the_oracle():
    jobs = WL.poisson_jobs(n_jobs=6, rate=20.0, seed=4, scale=0.03)
    assert {j["kind"] for j in jobs} == {"f32", "i16"} and {j["open_end"] for j in jobs} == {0, 1}
    print("serving: 6 Poisson jobs, 4 slots", flush=True)
    served = S.run_trace(jobs, "served", slots=4, max_rounds=40)
    for k, (j, g) in enumerate(zip(jobs, served["results"])):
        exp = O.perform_dba(j["seqs"], 0, j["open_end"], norm=1, max_rounds=40)
        assert g["medoid"] == exp["medoid"] and g["rounds"] == exp["rounds"], k
        assert np.array_equal(bits(g["centroid"]), bits(exp["centroid"])), k
    assert served["p99_s"] >= served["p50_s"] > 0
