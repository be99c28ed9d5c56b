"""Recorded DBA tracebacks (the cells writeDTWPath prints, io_utils.hpp:285-323)
and the openDBA command line end to end on the reference's KAT files.

The recorded cells are checked against a walk over the oracle's step matrix
of the same alignment (orientation flipped as dba.hpp:497-501 says), so the
path files carry the reference's cells and codes. The CLI run on the KAT
input files (tests/openDBA_test.cu) checks <prefix>.avg.txt against the
oracle's centroid as printed, and the format of the per-sequence path files.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from opendba_amd import native as N

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "opendba_amd", "lib", "openDBA")
MOVE = {1: (-1, -1), 2: (0, -1), 3: (-1, 0), 4: (0, -1)}
NAMES = {1: "DIAG", 2: "RIGHT", 3: "UP", 4: "OPEN_RIGHT", 255: "NIL", 5: "NIL_OPEN_RIGHT"}


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def oracle_walk(M):
    """Cells and codes from the top-right corner to the anchor."""
    i, j = M.shape[0] - 1, M.shape[1] - 1
    cells = []
    while True:
        code = int(M[i, j])
        cells.append((i, j, code))
        if code in (255, 5):
            return cells
        di, dj = MOVE[code]
        i, j = i + di, j + dj
        assert i >= 0 and j >= 0


def _walks(rng, lens):
    return [np.cumsum(rng.normal(size=int(L))).astype(np.float32) for L in lens]


@pytest.mark.parametrize("os_,oe", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_recorded_paths_match_oracle_walk(ctx, os_, oe):
    rng = np.random.default_rng(40 + 2 * os_ + oe)
    seqs = O.znorm(_walks(rng, [700, 1300, 2100, 900, 1024]))[0]
    cen = seqs[1].copy()  # length 1300: shorter than some members (flip with open_end)
    with N.DbaSession(ctx, seqs, os_, oe) as s:
        s.set_centroid(cen)
        s.record_paths(True)
        s.step()
        for k, q in enumerate(seqs):
            flip = oe and len(cen) < len(q)
            Y, X = (cen, q) if flip else (q, cen)
            _, M, _, _ = O.dtw_pair(Y, X, os_=os_, oe=oe, moves=True)
            exp = oracle_walk(M)
            i, j, code = s.path(k)
            print(f"mode os={os_} oe={oe} seq {k} len {len(q)} flip={flip} steps {len(i)}", flush=True)
            assert len(i) == len(exp)
            assert np.array_equal(i, [e[0] for e in exp])
            assert np.array_equal(j, [e[1] for e in exp])
            assert np.array_equal(code, [e[2] for e in exp])


def test_recorded_paths_stripe_equal_full_in_global_mode(ctx):
    rng = np.random.default_rng(51)
    seqs = O.znorm(_walks(rng, [1500, 2600, 900]))[0]
    cen = seqs[0].copy()
    paths = {}
    for pol in (N.PATH_FULL, N.PATH_STRIPE):
        with N.DbaSession(ctx, seqs, 0, 0, path_policy=pol) as s:
            s.set_centroid(cen)
            s.record_paths(True)
            s.step()
            paths[pol] = [s.path(k) for k in range(len(seqs))]
    for a, b in zip(paths[N.PATH_FULL], paths[N.PATH_STRIPE]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("oe", [0, 1])
def test_recorded_stripe_paths_reproduce_counts(ctx, oe):
    """Stripe mode in open_end has its own codes (the recompute height cut);
    the counted cells of the recorded paths must give the round's counts."""
    rng = np.random.default_rng(60 + oe)
    seqs = O.znorm(_walks(rng, [900, 2600, 1500, 3100, 700]))[0]
    cen = seqs[1].copy()
    with N.DbaSession(ctx, seqs, 0, oe, path_policy=N.PATH_STRIPE) as s:
        s.set_centroid(cen)
        s.record_paths(True)
        _, counts, _ = s.step()
        tally = np.zeros(len(cen), dtype=np.int64)
        for k, q in enumerate(seqs):
            flip = oe and len(cen) < len(q)
            i, j, code = s.path(k)
            c = i if flip else j
            counted = (code != 4) & (code != 5)
            np.add.at(tally, c[counted], 1)
    assert np.array_equal(tally, counts.astype(np.int64))


def _g6(v):
    return "%.6g" % v


def test_cli_kat_end_to_end(tmp_path):
    """openDBA on the reference's KAT files. The CLI always z-normalises: its
    -n flag ends up in argv[1] after GNU getopt's permutation and is then
    rejected as a read mode (openDBA.cu:19-28 with unistd.h getopt), as in
    the reference."""
    assert os.path.exists(CLI), "openDBA is not built"
    K = json.load(open(os.path.join(G, "kat_expected.json")))["different"]
    files = [os.path.join(G, "test2", f) for f in K["files"]]
    prefix = str(tmp_path / "kat_diff")
    print("cli: openDBA text float global (KAT 'different' files)", flush=True)
    r = subprocess.run([CLI, "text", "float", "global", prefix, "0", "/dev/null", "1.0"] + files,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    line = open(prefix + ".avg.txt").read().strip().split("\n")[-1].split("\t")
    vals = np.array([float(v) for v in line[1:]], dtype=np.float32)
    assert len(vals) == K["n"]
    seqs = [O.read_text_series(f, np.float32) for f in files]
    exp = O.perform_dba(seqs, norm=1)["centroid"]
    assert [_g6(v) for v in vals] == [_g6(v) for v in exp]
    # path files of the last round: name line, then 5 tab-separated fields
    order = sorted(range(len(seqs)), key=lambda k: len(seqs[k]))
    zs = O.znorm([seqs[k] for k in order])[0]
    for k, src in enumerate(order):
        lines = open(f"{prefix}.0.path{k}.txt").read().rstrip("\n").split("\n")
        name = files[src]  # cut at the first "." of the path (openDBA.cuh:58-59)
        assert lines[0] == (name[: name.index(".")] if "." in name else name)
        rows = [ln.split("\t") for ln in lines[1:]]
        assert all(len(r) == 5 for r in rows)
        assert rows[-1][4] in ("NIL", "NIL_OPEN_RIGHT")
        assert all(r[4] in NAMES.values() for r in rows)
        si = [int(r[0]) for r in rows]
        assert si[0] == len(seqs[src]) - 1 and si[-1] == 0
        assert all(a >= b for a, b in zip(si, si[1:]))
        assert all(r[1] == _g6(zs[k][int(r[0])]) for r in rows)
    r = subprocess.run([CLI, "-n", "text", "float", "global", prefix + "_n", "0", "/dev/null", "1.0"] + files,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "First argument (-n)" in r.stderr


def test_cli_clusters_end_to_end(tmp_path):
    """openDBA with cdist = 2 (two clusters of two or more members):
    <prefix>.cluster_membership.txt, one avg.txt line per cluster, each
    equal to the oracle's DBA loop for that cluster rescaled by its medoid,
    and the per-cluster path files."""
    from opendba_amd import native as N
    rng = np.random.default_rng(77)
    files, raw = [], []
    for g in range(2):
        base = np.cumsum(rng.normal(size=700)) * (1 + 2 * g)
        for r in range(3):
            L = int(rng.integers(400, 700))
            x = (base[:L] + rng.normal(0, 0.2, size=L)).astype(np.float32)
            f = tmp_path / f"g{g}r{r}.txt"
            f.write_text("\n".join("%.6g" % v for v in x) + "\n")
            files.append(str(f))
    prefix = str(tmp_path / "cl")
    print("cli: openDBA text float global cdist=2", flush=True)
    r = subprocess.run([CLI, "text", "float", "global", prefix, "0", "/dev/null", "2"] + files,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    seqs = [O.read_text_series(f, np.float32) for f in files]
    order = sorted(range(len(seqs)), key=lambda k: len(seqs[k]))
    names = [files[k][: files[k].index(".")] for k in order]
    z, mu, sg = O.znorm([seqs[k] for k in order])
    d = O.all_pairs(z, 0, 0)
    memb, med = N.cluster(d, [len(x) for x in z], 2.0)
    lines = open(prefix + ".cluster_membership.txt").read().rstrip("\n").split("\n")
    assert lines[0] == "## cluster distance threshold was 2"
    for i, ln in enumerate(lines[1:]):
        nm, c, mname = ln.split("\t")
        assert nm == names[i] and int(c) == memb[i] and mname == names[med[memb[i]]]
    avg = [ln.split("\t") for ln in open(prefix + ".avg.txt").read().rstrip("\n").split("\n")]
    assert len(avg) == len(med)
    for c, m in enumerate(med):
        members = [z[i] for i in range(len(z)) if memb[i] == c]
        cen = z[m].copy()
        if len(members) > 1:
            for _ in range(250):
                rc, new, _, delta = O.dba_update(cen, members, os_=0, oe=0)
                cen = new
                if delta == 0:
                    break
            for k in range(len(members)):
                assert os.path.exists(f"{prefix}.{c}.path{k}.txt")
        exp = (mu[m] + cen.astype(np.float64) * sg[m]).astype(np.float32)
        assert avg[c][0] == names[m]
        assert avg[c][1:] == [_g6(v) for v in exp], c
