"""The drop-in C++ layer and the openDBA command line against the oracle.

- tests/cpp/drop_in_calls.cpp, compiled against include/odba/opendba.hpp,
  calls every template entry point with the reference's argument lists:
  its setupAndRun runs the reference's KATs (tests/openDBA_test.cu:66-299),
  and what its other calls write is compared with the oracle.
- The command line: the two-stage 'minseg,minseg2' flow (cluster the
  segmented sequences, build the consensus on the unsegmented ones from
  the membership file; openDBA.cuh:121-133,180-184, dba.hpp:932-936), the
  resume rule and evolving-centroid checkpoint (dba.hpp:981-984,1046-1048),
  the free-memory stripe fallback (dba.hpp:515-548), the TSV and binary
  readers (cpu_utils.hpp:60-90,381-456) and pair_dists.txt
  (dba.hpp:188-216).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import dropin_expect as E
import oracle_lib as O
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "opendba_amd", "lib", "openDBA")
CALLS = os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_calls")


def run(cmd, env=None, timeout=120):
    print("run:", " ".join(os.path.basename(c) if os.sep in c else c for c in cmd[:8]), flush=True)
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=e)


def write_text(path, x):
    path.write_text("".join("%.9g\n" % v for v in x))
    return str(path)


def read_lines(path):
    return open(path).read().rstrip("\n").split("\n")


def walk_groups(rng, groups=2, per=3, lo=300, hi=700):
    out = []
    for g in range(groups):
        base = np.cumsum(rng.normal(size=hi)) * (1 + 2 * g)
        for _ in range(per):
            L = int(rng.integers(lo, hi))
            out.append((base[:L] + rng.normal(0, 0.2, size=L)).astype(np.float32))
    return out


# ---------------------------------------------------------------------------
# the compiled reference-style caller
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["same", "different", "three", "truncated", "large"])
def test_drop_in_setupAndRun_kats(tmp_path, name):
    K = json.load(open(os.path.join(G, "kat_expected.json")))[name]
    files = [os.path.join(G, "test2", f) for f in K["files"]]
    prefix = str(tmp_path / "kat")
    r = run([CALLS, "kat", prefix, str(K["open_end"])] + files)
    assert r.returncode == 0, r.stderr[-2000:]
    row = read_lines(prefix + ".avg.txt")[0].split("\t")
    vals = np.array([np.float32(v) for v in row[1:]])
    assert len(vals) == K["n"]
    for i, v in K["exact"].items():
        assert vals[int(i)] == np.float32(v)
    for i, v in K["round3"].items():
        assert np.float32(float("%.3f" % vals[int(i)])) == np.float32(v), (i, vals[int(i)], v)


def test_drop_in_every_template(tmp_path):
    rng = np.random.default_rng(123)
    raw = walk_groups(rng)
    files = [write_text(tmp_path / f"s{i}.txt", x) for i, x in enumerate(raw)]
    out = tmp_path / "out"
    out.mkdir()
    r = run([CALLS, "calls", str(out)] + files)
    assert r.returncode == 0, r.stderr[-3000:]
    n = len(raw)
    seqs = [O.read_text_series(f, np.float32) for f in files]
    z, mu, sg = O.znorm(seqs)
    # normalizeSequences, both overloads
    assert np.array_equal(np.fromfile(out / "znorm_mu.bin"), mu)
    assert np.array_equal(np.fromfile(out / "znorm_sigma.bin"), sg)
    for i in range(n):
        for tag in ("znorm_", "znorm5_"):
            assert np.array_equal(np.fromfile(out / f"{tag}{i}.bin", np.float32).view(np.uint32), z[i].view(np.uint32))
    # approximateMedoidIndices (file order, one cluster)
    d = O.all_pairs(z, 0, 0)
    m, _ = O.medoid_single(d, [len(x) for x in z])
    assert int(np.fromfile(out / "medoid.bin", np.int32)[0]) == m
    pd = read_lines(out / "medoid.pair_dists.txt")
    assert len(pd) == n
    first = pd[0].split("\t")
    assert first[0] == files[0] and first[1] == "0"  # readers keep the file name as given
    assert first[2:] == [E.g6(v) for v in d[: n - 1]]
    # DBAUpdate from the medoid
    rc, exp, _, delta = O.dba_update(z[m], z, 0, 0)
    assert rc == 0
    assert np.array_equal(np.fromfile(out / "dba_update.bin", np.float32).view(np.uint32), exp.view(np.uint32))
    assert np.fromfile(out / "dba_delta.bin")[0] == delta
    assert all(os.path.exists(out / f"round.path{k}.txt") for k in range(n))
    # adaptive_segmentation and writeSequences
    ev = O.kseg(z, 8, 0, 65536)
    for i in range(n):
        assert np.array_equal(np.fromfile(out / f"kseg_{i}.bin", np.float32).view(np.uint32), ev[i].view(np.uint32))
    segl = read_lines(out / "segmented_seqs.txt")
    assert [ln.split("\t")[1:] for ln in segl] == [[E.g6(v) for v in e] for e in ev]
    # chopPrefixFromSequences: the first 40 values of sequence 0 as prefix
    chop = [ln.split("\t") for ln in read_lines(out / "chop.prefix_chop.txt")]
    for i in range(n):
        c, h = O.prefix_chop(z[0][:40].copy(), z[i])
        assert int(chop[i][1]) == c and int(chop[i][2]) == len(z[i])
        assert [int(v) for v in chop[i][3:]] == h.tolist()
    # performDBA: CLUSTER_ONLY (cdist 2) then CONSENSUS_ONLY on the same
    # arrays (sorted and z-normed again, as the second call does)
    names = [E.stem(f) for f in files]
    s1, nm1, memb, med, _ = E.cluster_only(seqs, names, 0, 0, 1, 2.0)
    assert read_lines(out / "perf.cluster_membership.txt") == E.membership_lines(nm1, memb, med, 2.0)
    s2, mu2, sg2 = O.znorm(s1)
    assert read_lines(out / "perf.avg.txt") == E.consensus_lines(s2, nm1, memb, med, 0, 0, True, mu2, sg2)
    # readMedoidIndices on that file
    assert np.fromfile(out / "read_membership.bin", np.int32).tolist() == memb
    assert np.fromfile(out / "read_medoids.bin", np.int32).tolist() == med[:2]


# ---------------------------------------------------------------------------
# the command line
# ---------------------------------------------------------------------------

def _squiggle_files(tmp_path):
    """Six squiggles whose event counts do not follow their lengths, so the
    segmented and unsegmented length orders differ."""
    rng = np.random.default_rng(31)
    spec = [(4000, 40.0), (1500, 6.0), (3000, 30.0), (2000, 8.0), (3500, 12.0), (2500, 5.0)]
    raw = []
    for L, dwell in spec:
        raw.append(WL.squiggle(L, rng, dwell_mean=dwell).astype(np.float32))
    files = [write_text(tmp_path / f"q{i}.txt", x) for i, x in enumerate(raw)]
    return files, [O.read_text_series(f, np.float32) for f in files]


def test_cli_two_stage_cluster_segmented_consensus_raw(tmp_path):
    files, raw = _squiggle_files(tmp_path)
    names = [E.stem(f) for f in files]
    prefix = str(tmp_path / "two")
    r = run([CLI, "text", "float", "global", prefix, "8,0", "/dev/null", "2"] + files, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    seg = O.kseg(raw, 8, 0, 65536)
    assert all(len(s) >= 2 for s in seg)
    assert E.stable_order([len(s) for s in seg]) != E.stable_order([len(x) for x in raw])
    assert [ln.split("\t") for ln in read_lines(prefix + ".segmented_seqs.txt")] == \
        [[nm] + [E.g6(v) for v in s] for nm, s in zip(names, seg)]
    # stage 1: clusters of the segmented sequences; the names array is
    # sorted by segmented length in place
    _, nm1, memb1, med1, _ = E.cluster_only(seg, names, 0, 0, 1, 2.0)
    assert read_lines(prefix + ".cluster_membership.txt") == E.membership_lines(nm1, memb1, med1, 2.0)
    # stage 2: the unsegmented sequences (input order) with that names
    # array, both sorted by unsegmented length; memberships and medoids by name
    order = E.stable_order([len(x) for x in raw])
    nm2 = [nm1[k] for k in order]
    z, mu, sg = O.znorm([raw[k] for k in order])
    memb2 = [memb1[nm1.index(nm)] for nm in nm2]
    med2 = [nm2.index(nm1[m]) for m in med1]
    assert read_lines(prefix + ".avg.txt") == E.consensus_lines(z, nm2, memb2, med2, 0, 0, True, mu, sg)


def test_cli_two_stage_without_membership_file_exits_43(tmp_path):
    """With cdist 1 the clustering call writes no membership file
    (dba.hpp:951), so the consensus call cannot read one (io_utils.hpp:163-166)."""
    files, _ = _squiggle_files(tmp_path)
    r = run([CLI, "text", "float", "global", str(tmp_path / "nofile"), "8,0", "/dev/null", "1"] + files, timeout=300)
    assert r.returncode == 43, (r.returncode, r.stderr[-1000:])
    assert "Cannot open sequence cluster membership file" in r.stderr


def test_cli_resume_from_avg_and_checkpoint(tmp_path):
    rng = np.random.default_rng(77)
    raw = walk_groups(rng)
    files = [write_text(tmp_path / f"g{i}.txt", x) for i, x in enumerate(raw)]
    names = [E.stem(f) for f in files]
    seqs = [O.read_text_series(f, np.float32) for f in files]
    s, nm, memb, med, _ = E.cluster_only(seqs, names, 0, 0, 1, 2.0)
    z, mu, sg = O.znorm([seqs[k] for k in E.stable_order([len(x) for x in seqs])])
    assert max(memb) >= 1
    prefix = str(tmp_path / "res")
    seeded = "seeded\t1\t2"
    open(prefix + ".avg.txt", "w").write(seeded + "\n")
    # a partly converged centroid for cluster 1, as writeCentroidCheckpointToFile prints it
    start = (z[med[1]] * np.float32(0.9)).astype(np.float32)
    ck = f"{prefix}.1.evolving_centroid.txt"
    open(ck, "w").write(" ".join(E.g6(v) for v in start) + "\n")
    parsed = np.array([np.float32(t) for t in open(ck).read().split()], np.float32)
    r = run([CLI, "text", "float", "global", prefix, "0", "/dev/null", "2"] + files)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Restarting convergence with cluster 2/" in r.stderr
    assert "Resuming convergence from partially converged centroid" in r.stderr
    assert not os.path.exists(ck)
    exp = E.consensus_lines(z, nm, memb, med, 0, 0, True, mu, sg, first=1, checkpoints={1: parsed})
    assert read_lines(prefix + ".avg.txt") == [seeded] + exp


@pytest.mark.parametrize("mode,os_,oe", [("open_end", 0, 1), ("global", 0, 0)])
def test_cli_stripe_fallback_on_low_free_memory(tmp_path, mode, os_, oe):
    rng = np.random.default_rng(5 + oe)
    raw = [np.cumsum(rng.normal(size=int(rng.integers(300, 1500)))).astype(np.float32) for _ in range(4)]
    files = [write_text(tmp_path / f"w{i}.txt", x) for i, x in enumerate(raw)]
    seqs = [O.read_text_series(f, np.float32) for f in files]
    prefix = str(tmp_path / "st")
    r = run([CLI, "text", "float", mode, prefix, "0", "/dev/null", "1"] + files, env={"ODBA_REF_FREE_MEM": "1"})
    assert r.returncode == 0, r.stderr[-2000:]
    res = O.perform_dba(seqs, os_=os_, oe=oe, norm=1, path_policy=0, free_mem=1)
    name = E.stem(files[int(res["perm"][res["medoid"]])])
    assert read_lines(prefix + ".avg.txt") == ["\t".join([name] + [E.g6(v) for v in res["centroid"]])]


def test_cli_tsv_reader(tmp_path):
    rng = np.random.default_rng(9)
    raw = walk_groups(rng, groups=1, per=4, lo=200, hi=500)
    tsv = tmp_path / "set.tsv"
    tsv.write_text("".join("read%d\t" % i + "\t".join("%.9g" % v for v in x) + "\n" for i, x in enumerate(raw)))
    prefix = str(tmp_path / "tsv")
    r = run([CLI, "tsv", "float", "global", prefix, "0", "/dev/null", "1", str(tsv)])
    assert r.returncode == 0, r.stderr[-2000:]
    seqs = [np.array([np.float32(t) for t in ln.split("\t")[1:]], np.float32) for ln in read_lines(tsv)]
    res = O.perform_dba(seqs, norm=1)
    names = ["read%d" % int(res["perm"][k]) for k in range(len(seqs))]
    assert read_lines(prefix + ".avg.txt") == ["\t".join([names[res["medoid"]]] + [E.g6(v) for v in res["centroid"]])]
    # pair_dists.txt: the upper triangle row by row, names in sorted order
    pd = [ln.split("\t") for ln in read_lines(prefix + ".pair_dists.txt")]
    d = res["dists"]
    n, off = len(seqs), 0
    for rr in range(n - 1):
        assert pd[rr][0] == names[rr] and pd[rr][1 + rr] == "0"
        assert pd[rr][2 + rr:] == [E.g6(v) for v in d[off:off + n - rr - 1]]
        off += n - rr - 1


@pytest.mark.parametrize("ctype", ["short", "float", "double"])
def test_cli_binary_reader(tmp_path, ctype):
    rng = np.random.default_rng(11)
    files, seqs = [], []
    for i in range(4):
        L = int(rng.integers(200, 600))
        if ctype == "short":
            x = (90 + 12 * np.cumsum(rng.normal(size=L)) / np.sqrt(L)).astype(np.int16)
            seqs.append(x.astype(np.float32))  # int16 expanded to float (cpu_utils.hpp:81-85)
        else:
            x = np.cumsum(rng.normal(size=L)).astype(np.float32 if ctype == "float" else np.float64)
            seqs.append(x)
        f = tmp_path / f"b{i}.bin"
        x.tofile(f)
        files.append(str(f))
    prefix = str(tmp_path / "bin")
    r = run([CLI, "binary", ctype, "global", prefix, "0", "/dev/null", "1"] + files)
    assert r.returncode == 0, r.stderr[-2000:]
    res = O.perform_dba(seqs, norm=1)
    name = E.stem(files[int(res["perm"][res["medoid"]])])
    assert read_lines(prefix + ".avg.txt") == ["\t".join([name] + [E.g6(v) for v in res["centroid"]])]
