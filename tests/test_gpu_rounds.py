"""Packed DBA rounds across jobs (odba_rounds_*, the continuous batching of
the server): every job's result - centroid bits, medoid, rounds - equals
the plain per-job path (odba_perform_dba) and, for small jobs, the oracle,
whatever jobs it is packed with, at whichever step it joins and however the
step words were allocated (by the preparing side, or on joining)."""
import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _jobs(seed, n_jobs, dtype=np.float32, lmin=150, lmax=2600, nmin=3, nmax=9):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_jobs):
        n = int(rng.integers(nmin, nmax + 1))
        lens = rng.integers(lmin, lmax, size=n)
        if k % 2:
            seqs = [WL.squiggle(int(L), rng).astype(np.float64) for L in lens]
        else:
            seqs = [np.cumsum(rng.normal(size=int(L))) for L in lens]
        if np.dtype(dtype).kind == "i":
            seqs = [np.rint(s * 40.0).astype(dtype) for s in seqs]
        else:
            seqs = [s.astype(dtype) for s in seqs]
        out.append(seqs)
    return out


def _bits(a):
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def _same(a, b):
    assert a["medoid"] == b["medoid"] and a["rounds"] == b["rounds"], (a["medoid"], b["medoid"], a["rounds"],
                                                                       b["rounds"])
    assert np.array_equal(_bits(a["centroid"]), _bits(b["centroid"]))


def _run_packed(ctx, ectx, jobs, os_, oe, max_rounds, prealloc=False, join_every=2, norm=1):
    """Prepares every job on ctx, then adds them to one engine on ectx two
    steps apart (continuous batching) and steps until all have ended. With
    prealloc, every other job's step words are allocated before it joins."""
    dtype = jobs[0][0].dtype
    prepared = [N.Job(ctx, s, os_, oe, norm, max_rounds) for s in jobs]
    if prealloc:
        for p in prepared[::2]:
            assert p.words(allocate=True) > 0
    by_handle = {p.h.value: k for k, p in enumerate(prepared)}
    res = [None] * len(jobs)
    try:
        with N.Rounds(ectx, dtype, os_, oe) as R:
            waiting = list(range(len(jobs)))
            step, held = 0, 0
            while waiting or held:
                if waiting and step % join_every == 0:
                    R.add(prepared[waiting.pop(0)])
                fin, held = R.step()
                for h in fin:
                    k = by_handle[h]
                    res[k] = prepared[k].result()
                    assert res[k]["rc"] == 0
                step += 1
                assert step < 10000
    finally:
        for p in prepared:
            p.close()
    return res


@pytest.mark.parametrize("oe", [0, 1])
@pytest.mark.parametrize("prealloc", [False, True])
def test_packed_rounds_equal_plain_path(oe, prealloc):
    jobs = _jobs(seed=7 + oe, n_jobs=6)
    print(f"packed rounds: oe={oe} prealloc={prealloc}", flush=True)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, 0, oe, norm=1, max_rounds=25) for s in jobs]
        packed = _run_packed(ctx, ectx, jobs, 0, oe, 25, prealloc)
    for p, q in zip(plain, packed):
        _same(p, q)


@pytest.mark.parametrize("os_,oe", [(1, 0), (1, 1)])
def test_packed_rounds_open_start_modes(os_, oe):
    jobs = _jobs(seed=21 + oe, n_jobs=4)
    print(f"packed rounds: os={os_} oe={oe}", flush=True)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, os_, oe, norm=1, max_rounds=20) for s in jobs]
        packed = _run_packed(ctx, ectx, jobs, os_, oe, 20, False, join_every=1)
    for p, q in zip(plain, packed):
        _same(p, q)


@pytest.mark.parametrize("dtype", [np.float64, np.int32])
def test_packed_rounds_other_value_types(dtype):
    jobs = _jobs(seed=31, n_jobs=4, dtype=dtype, lmax=1500)
    print(f"packed rounds: dtype={np.dtype(dtype).name}", flush=True)
    norm = 0 if np.dtype(dtype).kind == "i" else 1
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, 0, 0, norm=norm, max_rounds=15) for s in jobs]
        packed = _run_packed(ctx, ectx, jobs, 0, 0, 15, False, norm=norm)
    for p, q in zip(plain, packed):
        _same(p, q)


@pytest.mark.parametrize("oe", [0, 1])
def test_packed_rounds_equal_oracle(oe):
    jobs = _jobs(seed=41 + oe, n_jobs=3, lmin=100, lmax=900, nmax=6)
    print(f"packed rounds vs oracle: oe={oe}", flush=True)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        packed = _run_packed(ctx, ectx, jobs, 0, oe, 12, False, join_every=1)
    for s, q in zip(jobs, packed):
        exp = O.perform_dba(s, 0, oe, norm=1, max_rounds=12)
        _same(dict(medoid=exp["medoid"], rounds=exp["rounds"], centroid=exp["centroid"]), q)


def test_packed_rounds_stop_rule_and_leaving_jobs():
    """Jobs with different round limits leave at different steps; the rest
    keep their results (max_rounds 1, 3 and 40)."""
    jobs = _jobs(seed=51, n_jobs=3)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, 0, 1, norm=1, max_rounds=m) for s, m in zip(jobs, (1, 3, 40))]
        prepared = [N.Job(ctx, s, 0, 1, 1, m) for s, m in zip(jobs, (1, 3, 40))]
        with N.Rounds(ectx, np.float32, 0, 1) as R:
            for p in prepared:
                R.add(p)
            ended = []
            held = len(jobs)
            while held:
                fin, held = R.step()
                ended.append([prepared.index(next(p for p in prepared if p.h.value == h)) for h in fin])
        res = [p.result() for p in prepared]
        for p in prepared:
            p.close()
    assert ended[0] == [0]
    for p, q in zip(plain, res):
        _same(p, q)


@pytest.mark.parametrize("oe", [0, 1])
def test_packed_rounds_long_jobs_take_swath_bands(oe):
    """Jobs whose pairs span more than 16 swaths (sequences of 17k-21k
    samples: flipped onto the columns in open end, a long medoid in global
    mode) run their path jobs as swath bands inside the packed step; results
    equal the plain path, with a short job packed beside them."""
    rng = np.random.default_rng(61 + oe)
    long_job = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(17000, 21000, size=4)]
    short_job = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(300, 2000, size=5)]
    mixed_job = long_job[:2] + short_job[:3]
    jobs = [long_job, short_job, mixed_job]
    print(f"packed rounds with long jobs oe={oe}", flush=True)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, 0, oe, norm=1, max_rounds=4) for s in jobs]
        packed = _run_packed(ctx, ectx, jobs, 0, oe, 4, False, join_every=1)
    for p, q in zip(plain, packed):
        _same(p, q)


@pytest.mark.parametrize("oe", [0, 1])
def test_packed_rounds_nine_to_sixteen_swaths(oe):
    """Jobs whose widest path matrices hold 9-16 swaths (sequences of
    9k-16k samples): in open end that group runs as swath bands with
    second-phase claims inside the packed step, in global mode as strips;
    results equal the plain path (strips), with a short job beside them."""
    rng = np.random.default_rng(71 + oe)
    mid_job = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(9000, 16000, size=4)]
    short_job = [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in rng.integers(300, 2000, size=5)]
    mixed_job = mid_job[:2] + short_job[:3]
    jobs = [mid_job, short_job, mixed_job]
    print(f"packed rounds with 9-16 swath jobs oe={oe}", flush=True)
    with N.Context(0) as ctx, N.Context(0) as ectx:
        plain = [ctx.perform_dba(s, 0, oe, norm=1, max_rounds=4) for s in jobs]
        packed = _run_packed(ctx, ectx, jobs, 0, oe, 4, False, join_every=1)
    for p, q in zip(plain, packed):
        _same(p, q)


def _swath_words(ly, os_, oe):
    """Words of one 1024-column swath: 64 lanes x the rows of its skewed main
    loop rounded up to a multiple of 64 (the last row of an open-end swath
    and the first of an open-start one are not stored)."""
    nr = (ly - 2 if (oe and ly >= 2) else ly - 1) - (1 if os_ else 0) + 1
    return (nr + 63) * 64 if nr > 0 else 0


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 1)])
def test_job_words_size(os_, oe):
    """odba_job_words: every member's path words (flipped onto the columns
    in open end when the centroid is shorter) plus 64 top-row words per
    swath, 4 bytes each; allocating keeps the size and a second call is a
    no-op. odba_job_shape: the centroid length and the widest member's
    swath count."""
    jobs = _jobs(seed=71 + oe, n_jobs=2, lmin=100, lmax=5000)
    with N.Context(0) as ctx:
        for s in jobs:
            lc = len(ctx.perform_dba(s, os_, oe, norm=1, max_rounds=1)["centroid"])
            want, kmax = 0, 0
            for x in s:
                L = len(x)
                flip = oe and lc < L
                ly, lx = (lc, L) if flip else (L, lc)
                K = (lx + 1023) // 1024
                kmax = max(kmax, K)
                want += K * (_swath_words(ly, os_, oe) + 64) * 4
            p = N.Job(ctx, s, os_, oe, 1, 5)
            try:
                assert p.shape() == (lc, kmax)
                assert p.words() == want
                assert p.words(allocate=True) == want
                assert p.words(allocate=True) == want
            finally:
                p.close()
