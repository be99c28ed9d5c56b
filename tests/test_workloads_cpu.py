"""The seeded workload generators produce the configurations the benchmarks
are quoted on: counts, length ranges, value types, modes and arrival rate.
CPU only."""
import numpy as np

from opendba_amd import workloads as WL


def test_random_walks_small_method():
    seqs = WL.random_walks(32, 200, 400, seed=0)
    assert len(seqs) == 32
    assert all(s.dtype == np.float32 and 200 <= len(s) <= 400 for s in seqs)
    steps = np.concatenate([np.diff(s.astype(np.float64)) for s in seqs])
    assert abs(steps.std() - 1.0) < 0.05  # step sd 1


def test_squiggles_batched_uneven():
    seqs = WL.squiggles(2000, 2000, 20000, seed=1)
    lens = np.array([len(s) for s in seqs])
    assert len(seqs) == 2000 and lens.min() >= 2000 and lens.max() <= 20000
    assert all(s.dtype == np.int16 for s in seqs)
    # log-uniform: the median sits near the geometric mean of the bounds
    assert abs(np.log(np.median(lens)) - 0.5 * np.log(2000 * 20000)) < 0.15
    sample = np.concatenate(seqs[:50]).astype(np.float64)
    assert abs(sample.mean() - 90.0) < 3.0  # levels ~N(90, 12) pA
    assert WL.pair_cells(lens) > 1.2e14


def test_squiggle_dwell_and_noise():
    rng = np.random.default_rng(3)
    s = WL.squiggle(200_000, rng, level_sd=0.0, noise_sd=0.0).astype(np.int64)
    # with no level spread and no noise every sample is the level itself
    assert np.all(s == 90)
    rng = np.random.default_rng(4)
    s = WL.squiggle(200_000, rng, level_sd=50.0, noise_sd=0.0).astype(np.int64)
    runs = np.diff(np.flatnonzero(np.diff(s) != 0))
    assert 7.0 < runs.mean() < 9.0  # geometric dwell, mean 8 samples


def test_serving_trace_full_size():
    jobs = WL.poisson_jobs(n_jobs=32, rate=20.0, seed=4, scale=1.0)
    assert len(jobs) == 32
    arrivals = np.array([j["arrival"] for j in jobs])
    assert np.all(np.diff(arrivals) > 0)
    assert 0.02 < arrivals[-1] / 32 < 0.1  # mean gap 1/20 s
    counts = [len(j["seqs"]) for j in jobs]
    assert min(counts) >= 16 and max(counts) <= 512
    lens = np.concatenate([[len(s) for s in j["seqs"]] for j in jobs])
    assert lens.min() >= 500 and lens.max() <= 50_000
    assert {j["kind"] for j in jobs} == {"f32", "i16"}
    assert {j["open_end"] for j in jobs} == {0, 1}
    assert all(s.dtype == np.float32 for j in jobs for s in j["seqs"])


def test_serving_trace_is_seeded():
    a = WL.poisson_jobs(n_jobs=3, rate=20.0, seed=4, scale=0.1)
    b = WL.poisson_jobs(n_jobs=3, rate=20.0, seed=4, scale=0.1)
    for x, y in zip(a, b):
        assert x["arrival"] == y["arrival"] and x["open_end"] == y["open_end"]
        assert all(np.array_equal(p, q) for p, q in zip(x["seqs"], y["seqs"]))


def test_squiggle_reads_segment_then_cluster():
    reads, label = WL.squiggle_reads(n_templates=10, reads_per=20, lmin=4000, lmax=40000, seed=5)
    assert len(reads) == 200 and sorted(np.bincount(label)) == [20] * 10
    assert all(r.dtype == np.int16 for r in reads)
    # reads of one template vary in length around it (own dwell times)
    by_t = [np.array([len(r) for r, t in zip(reads, label) if t == k]) for k in range(10)]
    assert all(v.std() > 0 for v in by_t)


def test_stable_length_sort():
    seqs = [np.zeros(n, np.float32) + i for i, n in enumerate([5, 3, 5, 1, 3])]
    out, order = WL.stable_length_sort(seqs)
    assert order == [3, 1, 4, 0, 2]  # equal lengths keep their input order
    assert [len(s) for s in out] == [1, 3, 3, 5, 5]
