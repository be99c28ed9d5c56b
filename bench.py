"""All-vs-all DTW benchmark on one MI355X.

Workload: 2,000 synthetic int16 nanopore-like squiggles (levels ~N(90,12) pA,
geometric dwell mean 8, noise sd 2), lengths log-uniform 2k-20k, expanded to
float (CLI type "short"), stably sorted by length, z-normalised, then the
open-end all-vs-all DTW of all 1,999,000 pairs in one persistent launch.
A step is one full all-vs-all over the resident set. After the timed steps,
the same reads go through perform_dba (open end) to convergence, and its
wall time and round count are reported beside the line's value.

Metric: GCUPS = sum over pairs of |Y|*|X| (the full unbanded matrix, also for
pairs whose open-end search stops early) / wall time of the step;
evaluated_gcups counts only the cells the kernel evaluates (an open-end pair
stops at its abrogated swath).

The default line also carries, under a wall-time budget, the other
configurations the metric is quoted on: the small method (32 random walks,
global, 496 pairs: kernel time, roofline fraction, pipeline with 3 DBA
rounds), the full-size serving trace (served vs plain per-job jobs/s, p50,
p99, identical results) and segment-then-cluster per-cluster DBA (all
clusters' rounds packed vs one cluster at a time).

Roofline: the swath kernel is bound by VALU issue, not by HBM or MFMA. The
algorithmic work is 4 VALU lane-ops per evaluated cell (sub, mul, min3, add;
fp-contract off), counted over the swaths the kernel actually evaluated (an
open-end pair stops at its abrogated swath). The peak is MI355X's published
FP32 vector rate, 157.3 TFLOP/s = 78.6e12 lane-ops/s (128 lane-ops per CU per
clock, an FMA counting as two flops). The same object carries the issue rate
the cell update's own instruction mix reaches in a probe kernel on the same
GPU, and the HBM traffic of one launch read from the committed FETCH_SIZE /
WRITE_SIZE counter summary for this workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

T_START = time.perf_counter()
METRIC = "all-vs-all DTW GCUPS + DBA jobs/s with p50/p99 latency vs plain per-job path"
VALU_OPS_PER_CELL = 4      # sub, mul, min3, add (fp-contract off)
FP32_VECTOR_TFLOPS = 157.3  # MI355X published FP32 vector peak; 2 flops per FMA lane-op
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_all_pairs_squiggles{n}_open_end_hbm_traffic.md")


def committed_traffic(n):
    """HBM bytes per all-pairs launch from the committed counter summary for
    this workload size, or None when there is none."""
    import re
    path = PMC_SUMMARY.format(n=n)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def evaluated_cells(lens_sorted, flags):
    """Cells the kernel evaluated: swaths up to the open-end stop, if any."""
    L = np.asarray(lens_sorted, dtype=np.int64)
    n = len(L)
    tot = 0.0
    k = 0
    for r in range(n - 1):
        ly = L[r]
        lx = L[r + 1:]
        f = flags[k:k + n - r - 1]
        ab = np.where(f & 1, f >> 8, -1)
        cols = np.where(ab >= 0, np.minimum(lx, ab * 1024), lx)
        tot += float(ly) * float(cols.sum())
        k += n - r - 1
    return tot


def host_cpu():
    """CPU model, CPUs on the host and CPUs this process may run on."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


def cpu_baseline(seqs, target_s=15.0, os_=0, oe=1):
    """Oracle all-vs-all on a seeded uniform sample of pairs, timed on this
    host's cores: one thread per CPU the process may run on, at most 16 (a
    one-GPU share of the GPU box's host)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n = len(seqs)
    npairs = n * (n - 1) // 2
    model, host_cpus, allowed = host_cpu()
    threads = max(1, min(allowed, 16))
    rng = np.random.default_rng(123)
    lens = np.array([len(s) for s in seqs], dtype=np.float64)

    starts = np.array([r * (2 * n - r - 1) // 2 for r in range(n)], dtype=np.int64)

    def cells_of(sub):
        r = np.searchsorted(starts, sub, side="right") - 1
        c = sub - starts[r] + r + 1
        return float((lens[r] * lens[c]).sum())

    size = 8 * threads
    sub = np.sort(rng.choice(npairs, size=size, replace=False))
    t0 = time.perf_counter()
    O.all_pairs(seqs, os_, oe, nthreads=threads, subset=sub)
    dt = time.perf_counter() - t0
    size2 = int(max(size, min(npairs, size * target_s / max(dt, 1e-3))))
    sub = np.sort(rng.choice(npairs, size=size2, replace=False))
    reps = max(1, int(target_s / max(dt * size2 / size, 1e-3))) if size2 == npairs else 1
    t0 = time.perf_counter()
    for _ in range(reps):
        O.all_pairs(seqs, os_, oe, nthreads=threads, subset=sub)
    dt = time.perf_counter() - t0
    gc = reps * cells_of(sub) / dt / 1e9
    mode = "open_end" if oe else "global"
    return dict(value=gc, unit="GCUPS", cores=threads, kind="port",
                cpu_model=model, host_cpus=host_cpus, cpus_allowed=allowed,
                build="g++ -O3 -march=x86-64-v3 -ffp-contract=off -fwrapv (oracle/Makefile)",
                sample=f"{size2} seeded uniform pairs of the same workload x {reps} ({reps * cells_of(sub):.3g} "
                       f"full-matrix cells, {mode}, oracle with {threads} threads, {dt:.1f} s)")


def small32_leg(args, N, WL, steps=20, warmup=3):
    """The plain small method: 32 float32 random walks (step sd 1), lengths
    uniform in [200, 400], z-normalised, global all-vs-all (496 pairs), then
    medoid and 3 DBA rounds. GCUPS from the median all-pairs kernel time."""
    raw = WL.random_walks(32, 200, 400, seed=0)
    with N.Context(0) as ctx:
        info = ctx.device_info()
        seqs, _, _ = ctx.znorm(raw)
        seqs, _ = WL.stable_length_sort(seqs)
        lens = [len(s) for s in seqs]
        nominal = WL.pair_cells(lens)
        sset = ctx.seqset(seqs)
        t_tune = time.perf_counter()
        sched, tune_ms, tried = sset.tune(0, 0, reps=5)
        t_tune = time.perf_counter() - t_tune
        for w in range(warmup):
            sset.all_pairs(0, 0, copy_out=False)
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            _, _, ms = sset.all_pairs(0, 0, copy_out=False)
            kms.append(ms)
        wall = time.perf_counter() - t0
        sset.close()
        tp, te = [], []
        for _ in range(5):
            t1 = time.perf_counter()
            ctx.perform_dba(raw, 0, 0, norm=1, max_rounds=3)
            tp.append(time.perf_counter() - t1)
        os.environ["ODBA_DBA_GRAPH"] = "0"
        for _ in range(5):
            t1 = time.perf_counter()
            ctx.perform_dba(raw, 0, 0, norm=1, max_rounds=3)
            te.append(time.perf_counter() - t1)
        del os.environ["ODBA_DBA_GRAPH"]
    kmed = float(np.median(kms))
    achieved = nominal * VALU_OPS_PER_CELL / (kmed / 1000.0) / 1e12
    peak = FP32_VECTOR_TFLOPS / 2.0
    return dict(seqs=seqs, info=info, nominal=nominal, kmed=kmed, wall=wall, steps=steps, warmup=warmup,
                sched=sched, tried=tried, t_tune=t_tune, achieved=achieved, peak=peak,
                gcups=nominal / (kmed / 1000.0) / 1e9,
                pipeline_ms=float(np.median(tp)) * 1000.0, pipeline_ms_eager=float(np.median(te)) * 1000.0)


SMALL32_WORKLOAD = ("plain small: 32 float32 random walks (step sd 1), lengths U[200,400], z-norm, global "
                    "all-vs-all (496 pairs), then medoid + 3 DBA rounds")


def pair_kernel_name(sched, os_=0, oe=0):
    """The all-vs-all kernel a schedule (odba_schedule as a dict) runs."""
    if sched.get("kind") == 2:
        return f"k_all_pairs_small<C={sched['band_c']}, B={sched['ring']}, os={os_}, oe={oe}>"
    if sched.get("kind") == 1:
        return f"k_all_pairs_bands<float, C={sched['band_c']}, os={os_}, oe={oe}>"
    return f"k_all_pairs<float, os={os_}, oe={oe}>"


def small32_summary(m):
    """The small method's numbers as one object of the default line."""
    return {"workload": SMALL32_WORKLOAD, "pairs": 496, "full_matrix_cells": m["nominal"],
            "gcups": round(m["gcups"], 2), "kernel_ms_median": round(m["kmed"], 4),
            "kernel": pair_kernel_name(m["sched"]), "schedule": m["sched"],
            "roofline_frac": round(m["achieved"] / m["peak"], 4),
            "pipeline_ms_median": round(m["pipeline_ms"], 3),
            "pipeline": "perform_dba: z-norm, sort, all-vs-all, medoid, 3 DBA rounds, rescale"}


def run_small32(args, N, WL):
    m = small32_leg(args, N, WL, steps=max(args.steps, 20), warmup=max(args.warmup, 3))
    line = {
        "metric": METRIC, "value": round(m["gcups"], 2), "unit": "GCUPS", "n_gpus": 1, "steps": m["steps"],
        "warmup": m["warmup"], "ms_per_step": round(m["wall"] * 1000.0 / m["steps"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": SMALL32_WORKLOAD,
                   "pairs": 496, "full_matrix_cells": m["nominal"], "mode": "global", "device": m["info"]["name"],
                   "kernel_ms_median": round(m["kmed"], 4),
                   "schedule": m["sched"], "tuned_from": m["tried"], "tune_s": round(m["t_tune"], 3),
                   "pipeline_ms_median": round(m["pipeline_ms"], 3),
                   "pipeline_ms_median_without_graphs": round(m["pipeline_ms_eager"], 3),
                   "pipeline": "perform_dba: z-norm, sort, all-vs-all, medoid, 3 DBA rounds (graph replays "
                               "from the second), rescale"},
        "roofline": {"bound": "valu", "achieved": round(m["achieved"], 3), "peak": round(m["peak"], 3),
                     "unit": "TOP/s", "frac": round(m["achieved"] / m["peak"], 4), "traffic": None,
                     "kernel": pair_kernel_name(m["sched"]), "kernel_ms": round(m["kmed"], 4),
                     "ops_per_cell": VALU_OPS_PER_CELL},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(m["seqs"], target_s=10.0, os_=0, oe=0)
    return line


def run_stripes(args, N, WL):
    """Few very long pairs: 8 float32 sequences of length 1M-4M (a shared
    random-walk motif, time-warped per sequence by a seeded monotone warp,
    plus N(0, 0.1) noise), global full unbanded all-vs-all on the 28 pairs
    with the banded strip schedule. The result is checked against the same
    run on the reversed sequence order (Y and X swapped in every pair): the
    global distance is symmetric under transposition, bit for bit."""
    n, lmin, lmax = args.long_n, args.long_min, args.long_max
    seqs = WL.warped_motif_set(n, lmin, lmax, seed=2)
    print(f"stripes: {n} sequences, {sum(len(s) for s in seqs):,} samples", file=sys.stderr, flush=True)
    lens = [len(s) for s in seqs]
    nominal = WL.pair_cells(lens)
    with N.Context(0) as ctx:
        info = ctx.device_info()
        sset = ctx.seqset(seqs)
        kms = []
        d = None
        t0 = time.perf_counter()
        for step in range(args.steps):
            d, _, ms = sset.all_pairs(0, 0, copy_out=True)
            kms.append(ms)
            print(f"step {step}: kernel {ms:.1f} ms", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        sset.close()
        rev = ctx.seqset(seqs[::-1])
        dr, _, rms = rev.all_pairs(0, 0, copy_out=True)
        rev.close()
        print(f"reversed order: kernel {rms:.1f} ms", file=sys.stderr, flush=True)
    # pair (r, c) of the reversed set is pair (n-1-c, n-1-r) of the forward one, transposed
    iu = np.triu_indices(n, 1)
    F = np.zeros((n, n), dtype=np.float32)
    F[iu] = d
    R = np.zeros((n, n), dtype=np.float32)
    R[iu] = dr
    back = R[::-1, ::-1].T[iu]
    same = bool(np.array_equal(back.view(np.uint32), d.view(np.uint32)))
    # the DP traceback of every sequence against the medoid: one DBA round in
    # stripe mode (checkpointed block recomputes), with the paths recorded
    D = F + F.T
    med = int(np.argmin((D.astype(np.float64) ** 2).sum(axis=1)))
    tb = {}
    if not args.no_traceback:
        os_, oe = {"global": (0, 0), "open_end": (0, 1), "open_start": (1, 0), "open": (1, 1)}[args.stripe_mode]
        lc = len(seqs[med])
        with N.Context(0) as ctx2, N.DbaSession(ctx2, seqs, os_, oe, path_policy=N.PATH_STRIPE) as s:
            s.set_centroid(seqs[med])
            s.record_paths(True)
            t1 = time.perf_counter()
            newc, counts, delta = s.step()
            tb_s = time.perf_counter() - t1
            ok, steps = True, 0
            for k in range(n):
                pi, pj, pc = s.path(k)
                steps += len(pi)
                # from the matrix's far corner (the centroid on the rows when
                # an open-end pair is flipped) to (0, 0), monotone
                flip = oe and lc < len(seqs[k])
                h, w = (lc, len(seqs[k])) if flip else (len(seqs[k]), lc)
                ok = ok and pi[0] == h - 1 and pj[0] == w - 1 and pi[-1] == 0 and pj[-1] == 0
                ok = ok and bool(np.all(np.diff(pi) <= 0) and np.all(np.diff(pj) <= 0))
                ok = ok and bool(np.all((np.diff(pi) < 0) | (np.diff(pj) < 0)))
            counted = int(counts.astype(np.int64).sum())
            # global mode counts every path cell; the open modes leave their
            # open-right moves out
            ok = ok and (counted == steps if (os_, oe) == (0, 0) else 0 < counted <= steps)
        print(f"traceback round ({args.stripe_mode}): {tb_s:.1f} s, {steps:,} path cells, {counted:,} counted, "
              f"checks {ok}", file=sys.stderr, flush=True)
        fwd = float(sum(len(x) for x in seqs)) * lc
        tb = {"mode": args.stripe_mode, "dba_round_stripe_s": round(tb_s, 2), "medoid": med, "path_cells": steps,
              "counted_cells": counted, "forward_cells": fwd, "paths_monotone_and_counts_match": ok}
    kmed = float(np.median(kms))
    value = nominal / (kmed / 1000.0) / 1e9
    achieved = nominal * VALU_OPS_PER_CELL / (kmed / 1000.0) / 1e12
    peak = FP32_VECTOR_TFLOPS / 2.0
    return {
        "metric": METRIC, "value": round(value, 1), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(wall * 1000.0 / args.steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"few long pairs: {n} float32 warped-motif sequences, lengths U[{lmin},{lmax}], "
                               f"global full unbanded all-vs-all ({n * (n - 1) // 2} pairs), banded strip schedule",
                   "lengths": lens, "full_matrix_cells": nominal, "device": info["name"],
                   "kernel_ms_median": round(kmed, 1), "transposed_run_bit_identical": same,
                   "traceback": tb or "skipped"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 3), "unit": "TOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "k_all_pairs_bands<float, C=16, os=0, oe=0>", "kernel_ms": round(kmed, 1),
                     "ops_per_cell": VALU_OPS_PER_CELL},
    }


def serve_leg(args, N, WL, jobs_n=32, scale=1.0, slots=4, max_rounds=250):
    """Serving: a seeded Poisson trace (mean 20 jobs/s) of DBA jobs with
    16-512 series (scaled by scale), lengths log-uniform 500-50k (scaled),
    float32 random walks or int16 squiggles as float, global or open-end;
    each job is z-norm -> all-vs-all -> medoid -> DBA to convergence. The same
    trace runs through the plain per-job path (one job after another) and
    through the serving loop; jobs/s = jobs / makespan, latency = completion
    - arrival."""
    from opendba_amd import serve as S
    jobs = WL.poisson_jobs(n_jobs=jobs_n, rate=20.0, seed=4, scale=scale)
    cells = sum(WL.pair_cells([len(s) for s in j["seqs"]]) for j in jobs)
    log = lambda tag: (lambda k, o: print(f"{tag} job {k}: latency {o['latency']:.3f} s rounds {o['rounds']}",
                                          file=sys.stderr, flush=True))
    plain = S.run_trace(jobs, "plain", max_rounds=max_rounds, progress=log("plain"))
    served = S.run_trace(jobs, "served", slots=slots, max_rounds=max_rounds, progress=log("served"))
    same = all(a["centroid"].tobytes() == b["centroid"].tobytes() and a["medoid"] == b["medoid"]
               for a, b in zip(plain["results"], served["results"]))
    return {
        "workload": f"serving: {jobs_n} Poisson jobs (mean 20/s, seed 4), 16-512 series and lengths 500-50k scaled "
                    f"by {scale}, f32 walks / int16 squiggles, global / open_end, z-norm -> all-vs-all -> medoid -> "
                    f"DBA (max {max_rounds} rounds)",
        "slots": slots, "all_pairs_cells": cells,
        "served": {"jobs_per_s": round(served["jobs_per_s"], 4), "p50_s": round(served["p50_s"], 4),
                   "p99_s": round(served["p99_s"], 4), "makespan_s": round(served["makespan_s"], 3)},
        "plain": {"jobs_per_s": round(plain["jobs_per_s"], 4), "p50_s": round(plain["p50_s"], 4),
                  "p99_s": round(plain["p99_s"], 4), "makespan_s": round(plain["makespan_s"], 3)},
        "speedup_jobs_per_s": round(served["jobs_per_s"] / plain["jobs_per_s"], 3),
        "results_identical": bool(same),
    }


def run_serve(args, N, WL):
    leg = serve_leg(args, N, WL, jobs_n=args.jobs, scale=args.scale, slots=args.slots, max_rounds=args.max_rounds)
    return {
        "metric": METRIC, "value": leg["served"]["jobs_per_s"], "unit": "jobs/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(leg["served"]["makespan_s"] * 1000.0, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": leg["workload"], "slots": leg["slots"], "all_pairs_cells": leg["all_pairs_cells"]},
        "served": leg["served"], "plain": leg["plain"], "speedup_jobs_per_s": leg["speedup_jobs_per_s"],
        "results_identical": leg["results_identical"],
    }


def clusters_leg(args, N, WL):
    """Segment-then-cluster: int16 squiggles (levels N(90,12), dwell geom mean
    8, noise sd 2), lengths log-uniform 4k-40k, as reads of K molecules
    (n/K reads each, every read with its own dwell times and noise),
    k-segmentation (minimum segment length 8) to event medians, z-norm,
    all-vs-all on the event series, complete-linkage clustering into
    K clusters of two or more members, then DBA to convergence for every
    cluster: once with the plain per-cluster path (one DBA session after
    another) and once with all clusters' rounds packed into shared
    launches. Mode "open" (open start and open end) throughout: on these
    event series the global distances cluster by length alone."""
    reads_per = max(2, args.n_series // args.clusters)
    raw16, _ = WL.squiggle_reads(args.clusters, reads_per, 4000, 40000, seed=5)
    raw = WL.short_to_float(raw16)
    n = len(raw)
    samples = sum(len(s) for s in raw)
    stage = {}
    with N.Context(0) as ctx:
        t0 = time.perf_counter()
        ev = ctx.kseg(raw, 8)
        stage["kseg_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ev, mu, sg = ctx.znorm(ev)
        order = sorted(range(n), key=lambda i: len(ev[i]))
        ev = [ev[i] for i in order]
        lens = np.array([len(e) for e in ev], dtype=np.uint64)
        sset = ctx.seqset(ev)
        d, _, kms = sset.all_pairs(1, 1)
        sset.close()
        stage["allpairs_s"] = time.perf_counter() - t0
        cells = WL.pair_cells(lens)
        t0 = time.perf_counter()
        memb, med = N.cluster(d, lens, float(args.clusters))
        stage["cluster_s"] = time.perf_counter() - t0
        K = len(med)
        multi = [k for k in range(K) if (memb == k).sum() > 1]
        sizes = np.bincount(memb)
        print(f"clusters: {K} ({len(multi)} with 2+ members, largest {sizes.max()}), events {int(lens.sum())}, "
              f"all-pairs {kms:.1f} ms", file=sys.stderr, flush=True)
        R = args.max_rounds
        t0 = time.perf_counter()
        plain = {}
        for k in multi:  # the plain path: one cluster after another (dba.hpp:990-1110)
            members = [ev[i] for i in np.nonzero(memb == k)[0]]
            with N.DbaSession(ctx, members, 1, 1) as ses:
                ses.set_centroid(ev[med[k]])
                prev = np.zeros(len(ev[med[k]]), np.float32)
                prev2 = prev.copy()
                for r in range(R):
                    new, _, delta = ses.step()
                    rounds_k = r + 1
                    if delta == 0:
                        break
                    if r >= 1 and new.tobytes() == prev2.tobytes():
                        break
                    if r > 1:
                        prev2 = prev.copy()
                    if r > 0:
                        prev = new.copy()
                plain[k] = (new, rounds_k)
        plain_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        cents, rounds = ctx.dba_clusters(ev, memb, med, 1, 1, max_rounds=R)
        packed_s = time.perf_counter() - t0
    # bitwise: a centroid position no member maps to is NaN (0/0, dba.hpp:820-822)
    same = all(plain[k][0].tobytes() == cents[k].tobytes() and plain[k][1] == rounds[k] for k in multi)
    return {
        "workload": f"segment-then-cluster: {n} int16 squiggles as {reads_per} reads of each of {args.clusters} "
                    f"molecules, lengths 4k-40k ({samples} samples), k-seg minseg 8, z-norm, open (open start + "
                    f"open end) all-vs-all on events, complete linkage into K={args.clusters} clusters of 2+ "
                    f"members, per-cluster DBA (max {R} rounds)",
        "events": int(lens.sum()), "pairs": n * (n - 1) // 2, "all_pairs_cells": cells,
        "all_pairs_gcups": round(cells / (kms / 1000.0) / 1e9, 1), "clusters": K,
        "multi_member_clusters": len(multi), "rounds_total": int(sum(rounds[k] for k in multi)),
        **{k: round(v, 3) for k, v in stage.items()},
        "packed": {"jobs_per_s": round(len(multi) / packed_s, 3), "seconds": round(packed_s, 3)},
        "plain": {"jobs_per_s": round(len(multi) / plain_s, 3), "seconds": round(plain_s, 3)},
        "speedup_jobs_per_s": round(plain_s / packed_s, 3),
        "results_identical": bool(same),
    }


def run_clusters(args, N, WL):
    leg = clusters_leg(args, N, WL)
    return {
        "metric": METRIC, "value": leg["packed"]["jobs_per_s"], "unit": "jobs/s (cluster DBA)", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(leg["packed"]["seconds"] * 1000.0, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {k: v for k, v in leg.items() if k not in ("packed", "plain", "speedup_jobs_per_s",
                                                             "results_identical")},
        "packed": leg["packed"], "plain": leg["plain"], "speedup_jobs_per_s": leg["speedup_jobs_per_s"],
        "results_identical": leg["results_identical"],
    }


def run_dba(args, N, WL):
    """DBA rounds of the batched-uneven set: 2,000 int16 squiggles as float
    (lengths log-uniform 2k-20k), perform_dba in the given mode with
    --dba-rounds rounds at most (sanitize, sort, z-norm, all-vs-all, medoid,
    rounds, rescale). In global mode every round evaluates
    sum_i |S_i| * |centroid| cells (no open-end stop, no flip), which the line
    reports per round for the path kernel's roofline."""
    oe = 1 if args.mode == "open_end" else 0
    raw = WL.squiggles(args.n, 2000, 20000, seed=1)
    seqs = WL.short_to_float(raw)
    with N.Context(0) as ctx:
        info = ctx.device_info()
        t0 = time.perf_counter()
        r = ctx.perform_dba(seqs, 0, oe, norm=1, max_rounds=args.dba_rounds)
        wall = time.perf_counter() - t0
    lc = len(r["centroid"])
    cells_round = float(sum(len(s) for s in seqs)) * lc
    return {
        "metric": METRIC, "value": round(r["rounds"] / wall, 3), "unit": "DBA rounds/s (incl. all-vs-all)",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(wall * 1000.0, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"DBA rounds: {args.n} int16 squiggles as float, lengths 2k-20k, {args.mode}, "
                               f"perform_dba with at most {args.dba_rounds} rounds",
                   "mode": args.mode, "rounds": r["rounds"], "medoid": int(r["medoid"]), "centroid_len": lc,
                   "path_cells_per_round": cells_round if not oe else None,
                   "path_cells_per_round_upper": cells_round, "device": info["name"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--workload", choices=["squiggles", "small32", "serve", "clusters", "stripes", "dba"],
                    default="squiggles")
    ap.add_argument("--mode", choices=["global", "open_end"], default="open_end", help="dba: alignment mode")
    ap.add_argument("--long-n", type=int, default=8, help="stripes: sequences")
    ap.add_argument("--long-min", type=int, default=1_000_000, help="stripes: shortest length")
    ap.add_argument("--long-max", type=int, default=4_000_000, help="stripes: longest length")
    ap.add_argument("--no-traceback", action="store_true", help="stripes: skip the DBA round's tracebacks")
    ap.add_argument("--stripe-mode", choices=["global", "open_end", "open_start", "open"], default="global",
                    help="stripes: mode of the DBA round in stripe mode")
    ap.add_argument("--n-series", type=int, default=5000, help="clusters: squiggles")
    ap.add_argument("--clusters", type=int, default=250, help="clusters: K (clusters of 2+ members)")
    ap.add_argument("--jobs", type=int, default=32, help="serve: jobs in the trace")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="serve: series-count and length scale (1.0: the serving configuration as named)")
    ap.add_argument("--slots", type=int, default=4,
                    help="serve: preparation slots; the rounds of prepared jobs are packed by the server's engines")
    ap.add_argument("--max-rounds", type=int, default=250, help="serve: DBA round limit")
    ap.add_argument("--dba-rounds", type=int, default=250,
                    help="squiggles: DBA round limit of the perform_dba run after the timed all-vs-all (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the schedule autotuner")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="default run: wall-time budget; a leg runs only if its estimate still fits")
    ap.add_argument("--no-legs", action="store_true",
                    help="default run: only the headline all-vs-all (no small-method, serving or cluster legs)")
    ap.add_argument("--order", type=int, choices=[0, 1], default=None,
                    help="pin the pair packing order instead of tuning (0 built-in, 1 largest cells first)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the all-vs-all shards by "
              f"pair cost across GPUs by design, but this build measures one GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("serve", "squiggles"):
        # One hardware queue per slot stream of the serving loop (HIP reads
        # this when it starts; with fewer queues than streams, slots share
        # queues and their kernels run one after another). The default run
        # serves too, after its headline.
        want = min(32, args.slots + 20)  # the slots' streams and up to four rounds engines' five streams each
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < want:
            os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    from opendba_amd import native as N
    from opendba_amd import workloads as WL

    if args.workload == "small32":
        print(json.dumps(run_small32(args, N, WL)))
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args, N, WL)))
        return
    if args.workload == "clusters":
        print(json.dumps(run_clusters(args, N, WL)))
        return
    if args.workload == "stripes":
        print(json.dumps(run_stripes(args, N, WL)))
        return
    if args.workload == "dba":
        print(json.dumps(run_dba(args, N, WL)))
        return
    raw = WL.squiggles(args.n, 2000, 20000, seed=1)
    seqs, _ = WL.stable_length_sort(WL.short_to_float(raw))
    lens = [len(s) for s in seqs]
    nominal = WL.pair_cells(lens)
    with N.Context(0) as ctx:
        info = ctx.device_info()
        seqs, _, _ = ctx.znorm(seqs)
        sset = ctx.seqset(seqs)
        sched, tune_s, tried = None, 0.0, 0
        if args.order is not None:  # a pinned packing order (profiling runs)
            sched = dict(kind=0, order=args.order)
            sset.set_schedule(sched, 0, 1)
        elif not args.no_tune:  # pick the schedule (here: the pair packing order) on the full workload
            t_tune = time.perf_counter()
            sched, tune_ms, tried = sset.tune(0, 1, reps=1)
            tune_s = time.perf_counter() - t_tune
            print(f"tuned: {sched} ({tune_ms:.1f} ms) over {tried} candidates in {tune_s:.1f} s", file=sys.stderr,
                  flush=True)
        flags = None
        for w in range(args.warmup):
            _, flags, wms = sset.all_pairs(0, 1, copy_out=False, flags=(w == 0))
            print(f"warmup {w}: kernel {wms:.1f} ms", file=sys.stderr, flush=True)
        if flags is None:
            _, flags, _ = sset.all_pairs(0, 1, copy_out=False, flags=True)
        kms = []
        t0 = time.perf_counter()
        for s in range(args.steps):
            _, _, ms = sset.all_pairs(0, 1, copy_out=False)
            kms.append(ms)
            print(f"step {s}: kernel {ms:.1f} ms", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        sset.close()
        dba = None
        if args.dba_rounds > 0:  # the workload's second half: DBA to convergence on the same reads
            t1 = time.perf_counter()
            r = ctx.perform_dba(WL.short_to_float(raw), 0, 1, norm=1, max_rounds=args.dba_rounds)
            dwall = time.perf_counter() - t1
            dba = {"pipeline": "perform_dba open_end: sanitize, stable sort, z-norm, all-vs-all, medoid, DBA rounds "
                               f"until the stop rule or {args.dba_rounds} rounds, rescale",
                   "rounds": r["rounds"], "medoid": int(r["medoid"]), "centroid_len": int(len(r["centroid"])),
                   "wall_s": round(dwall, 3), "jobs_per_s": round(1.0 / dwall, 5),
                   "ms_per_round_est": round((dwall * 1000.0 - float(np.mean(kms))) / max(r["rounds"], 1), 2),
                   "ms_per_round_basis": "(wall - one all-vs-all kernel time) / rounds"}
            print(f"dba: {r['rounds']} rounds in {dwall:.1f} s", file=sys.stderr, flush=True)
    ms_per_step = wall * 1000.0 / args.steps
    value = nominal / (ms_per_step / 1000.0) / 1e9
    kavg = float(np.mean(kms))
    ev_cells = evaluated_cells(lens, flags)
    evaluated_gcups = ev_cells / (ms_per_step / 1000.0) / 1e9
    achieved = ev_cells * VALU_OPS_PER_CELL / (kavg / 1000.0) / 1e12
    peak = FP32_VECTOR_TFLOPS / 2.0
    from opendba_amd import probe
    mix_rate, _ = probe.cell_mix_rate(16)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GCUPS",
        "evaluated_gcups": round(evaluated_gcups, 2),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched uneven: {args.n} int16 squiggles (levels N(90,12), dwell geom mean 8, noise sd 2), "
                        f"lengths log-uniform 2k-20k, short->float, z-norm, open_end all-vs-all "
                        f"({args.n * (args.n - 1) // 2} pairs) in one persistent launch",
            "pairs": args.n * (args.n - 1) // 2,
            "full_matrix_cells": nominal,
            "evaluated_cells": ev_cells,
            "mode": "open_end",
            "device": info["name"],
            "schedule": sched if sched is not None else "built-in",
            "tuned_from": tried,
            "tune_s": round(tune_s, 1),
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(peak, 3),
            "unit": "TOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": committed_traffic(args.n),
            "kernel": "k_all_pairs<float, os=0, oe=1>" + (" (largest-cell-count-first packing)"
                                                           if sched and sched.get("order") == 1 else ""),
            "kernel_ms": round(kavg, 3),
            "ops_per_cell": VALU_OPS_PER_CELL,
            "peak_basis": "published FP32 vector 157.3 TFLOP/s / 2 flops per lane-op",
            "cell_mix_probe": round(mix_rate / 1e12, 3),
            "frac_of_probe": round(achieved / (mix_rate / 1e12), 4),
            "device_cus": info["compute_units"],
            "device_clock_mhz": info["clock_khz"] / 1000.0,
        },
    }
    if dba is not None:
        line["dba_to_convergence"] = dba
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(seqs)
    if not args.no_legs:
        legs_run(args, N, WL, line)
    print(json.dumps(line))


def legs_run(args, N, WL, line):
    """The other configurations the metric is quoted on, each as one object
    of the default line, in order, each only if its estimated wall time still
    fits the budget: the small method (its kernel, roofline fraction and
    pipeline), the full-size serving trace (served vs plain jobs/s, p50,
    p99, identical results) and the segment-then-cluster per-cluster DBA
    (packed vs one cluster at a time)."""
    legs = [("config1_small", 20.0, lambda: small32_summary(small32_leg(args, N, WL))),
            ("serving", 240.0, lambda: serve_leg(args, N, WL, jobs_n=32, scale=1.0, slots=args.slots)),
            ("segment_then_cluster", 30.0, lambda: clusters_leg(args, N, WL))]
    skipped = []
    for name, est, fn in legs:
        used = time.perf_counter() - T_START
        if used + est > args.budget_s:
            skipped.append(f"{name} (estimated {est:.0f} s, {used:.0f} s of the {args.budget_s:.0f} s budget used)")
            continue
        t0 = time.perf_counter()
        print(f"leg {name}: starting at {used:.0f} s", file=sys.stderr, flush=True)
        line[name] = fn()
        line[name]["leg_wall_s"] = round(time.perf_counter() - t0, 1)
    if skipped:
        line["legs_skipped"] = skipped


if __name__ == "__main__":
    main()
