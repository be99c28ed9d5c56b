"""Seeded synthetic inputs for the five workloads the hot path is measured on.

No datasets are downloaded; every generator is deterministic in its seed.
"""
import numpy as np


def random_walks(n=32, lmin=200, lmax=400, seed=0):
    """Gaussian random walks x_t = x_{t-1} + N(0, 1) in float32, lengths uniform in [lmin, lmax]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        out.append(np.cumsum(rng.normal(0.0, 1.0, size=L)).astype(np.float32))
    return out


def squiggle(L, rng, level_mu=90.0, level_sd=12.0, dwell_mean=8.0, noise_sd=2.0):
    """Nanopore-like current trace: piecewise-constant levels ~N(level_mu, level_sd) pA
    with geometric dwell times (mean dwell_mean samples) plus Gaussian noise, as int16."""
    nseg = int(L / dwell_mean * 1.5) + 8
    dw = rng.geometric(1.0 / dwell_mean, size=nseg)
    while dw.sum() < L:
        dw = np.concatenate([dw, rng.geometric(1.0 / dwell_mean, size=nseg)])
    lv = rng.normal(level_mu, level_sd, size=len(dw))
    sig = np.repeat(lv, dw)[:L] + rng.normal(0.0, noise_sd, size=L)
    return np.clip(np.rint(sig), -32768, 32767).astype(np.int16)


def squiggle_reads(n_templates=250, reads_per=20, lmin=4000, lmax=40000, seed=5, level_mu=90.0, level_sd=12.0,
                   dwell_mean=8.0, noise_sd=2.0):
    """Reads of shared molecules, as nanopore data has them: each template is
    a sequence of current levels ~N(level_mu, level_sd) pA sized for a
    log-uniform length in [lmin, lmax]; each read of it walks the same levels
    with its own geometric dwell times (mean dwell_mean samples) and
    Gaussian noise, as int16. Returns (reads, template index per read), reads
    in a seeded shuffled order."""
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(n_templates, lmin, lmax, rng)
    reads, label = [], []
    for t, L in enumerate(lens):
        lv = rng.normal(level_mu, level_sd, size=int(L / dwell_mean) + 1)
        for _ in range(reads_per):
            dw = rng.geometric(1.0 / dwell_mean, size=len(lv))
            sig = np.repeat(lv, dw) + rng.normal(0.0, noise_sd, size=int(dw.sum()))
            reads.append(np.clip(np.rint(sig), -32768, 32767).astype(np.int16))
            label.append(t)
    perm = rng.permutation(len(reads))
    return [reads[i] for i in perm], np.asarray(label)[perm]


def log_uniform_lengths(n, lmin, lmax, rng):
    return np.rint(np.exp(rng.uniform(np.log(lmin), np.log(lmax), size=n))).astype(np.int64)


def squiggles(n=2000, lmin=2000, lmax=20000, seed=1):
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(n, lmin, lmax, rng)
    return [squiggle(int(L), rng) for L in lens]


def warped_motif_set(n=8, lmin=1_000_000, lmax=4_000_000, seed=2, motif_len=None):
    """Shared random-walk motif, time-warped per sequence by a seeded monotone warp
    (cumulative sum of positive rates), plus N(0, 0.1) noise; float32."""
    rng = np.random.default_rng(seed)
    mlen = motif_len or lmax
    motif = np.cumsum(rng.normal(0.0, 1.0, size=mlen))
    out = []
    for _ in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        rates = rng.gamma(4.0, 0.25, size=L)  # positive, mean 1
        pos = np.cumsum(rates)
        pos = (pos - pos[0]) / (pos[-1] - pos[0]) * (mlen - 1)
        v = np.interp(pos, np.arange(mlen), motif) + rng.normal(0.0, 0.1, size=L)
        out.append(v.astype(np.float32))
    return out


def short_to_float(seqs):
    """CLI type "short": int16 input expanded to float (cpu_utils.hpp:81-85)."""
    return [s.astype(np.float32) for s in seqs]


def stable_length_sort(seqs):
    """Stable ascending length sort applied by performDBA (dba.hpp:876-885)."""
    order = sorted(range(len(seqs)), key=lambda i: len(seqs[i]))
    return [seqs[i] for i in order], order


def pair_cells(lens):
    """Full unbanded cell count of the all-vs-all: sum over pairs r<c of L_r * L_c."""
    lens = np.asarray(lens, dtype=np.float64)
    s = lens.sum()
    return float((s * s - (lens * lens).sum()) / 2.0)


def poisson_jobs(n_jobs=64, rate=20.0, seed=4, nmin=16, nmax=512, lmin=500, lmax=50_000, scale=1.0):
    """Seeded Poisson job trace: arrival times, and per job N series of
    log-uniform lengths, dtype float32 or int16->float, mode global or open_end.
    `scale` shrinks series counts and lengths for reduced traces."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    jobs = []
    for k in range(n_jobs):
        N = max(2, int(rng.integers(nmin, nmax + 1) * scale))
        lens = log_uniform_lengths(N, max(8, int(lmin * scale)), max(16, int(lmax * scale)), rng)
        kind = "f32" if rng.random() < 0.5 else "i16"
        mode = "global" if rng.random() < 0.5 else "open_end"
        if kind == "f32":
            seqs = [np.cumsum(rng.normal(0.0, 1.0, size=int(L))).astype(np.float32) for L in lens]
        else:
            seqs = [squiggle(int(L), rng).astype(np.float32) for L in lens]
        jobs.append(dict(arrival=float(t[k]), seqs=seqs, kind=kind, open_end=int(mode == "open_end")))
    return jobs
