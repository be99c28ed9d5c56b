"""Prefix location (chopPrefixFromSequences, dba.hpp:1192-1395) on the GPU
against the oracle: chop positions and leader-path histograms, bit-exact,
with and without z-normalisation, with sequences long enough for the
open-end early stop and shorter than the prefix."""
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


def _set(rng, n, P, dtype=np.float32):
    base = np.cumsum(rng.normal(size=P + 6000))
    prefix = base[:P] + rng.normal(scale=0.05, size=P)
    seqs = []
    for k in range(n):
        L = int(rng.integers(20, 5000)) if k % 5 else int(rng.integers(1100, 4000))
        head = base[: min(L, P + 40)] + rng.normal(scale=0.2, size=min(L, P + 40))
        tail = np.cumsum(rng.normal(size=max(0, L - len(head)))) + head[-1]
        seqs.append(np.concatenate([head, tail])[:L])
    cast = (lambda a: np.rint(a * 10).astype(dtype)) if np.dtype(dtype).kind == "i" else (lambda a: a.astype(dtype))
    return cast(prefix), [cast(s) for s in seqs]


def _oracle(prefix, seqs, norm):
    if norm:
        seqs = O.znorm(seqs)[0]
        prefix = O.znorm([prefix])[0][0]
    out = [O.prefix_chop(prefix, s) for s in seqs]
    return np.array([o[0] for o in out], dtype=np.uint64), np.stack([o[1] for o in out]), seqs


@pytest.mark.parametrize("dtype,norm", [(np.float32, 1), (np.float32, 0), (np.float64, 1), (np.int32, 0)])
def test_prefix_chop_matches_oracle(ctx, dtype, norm):
    rng = np.random.default_rng(40 + norm)
    prefix, seqs = _set(rng, 24, 80, dtype)
    print(f"prefix chop {np.dtype(dtype).name} norm={norm}", flush=True)
    chop, hist, got_seqs = ctx.prefix_chop(prefix, seqs, norm=norm)
    echop, ehist, eseqs = _oracle(prefix, seqs, norm)
    assert np.array_equal(chop, echop)
    assert np.array_equal(hist, ehist)
    for g, e in zip(got_seqs, eseqs):
        assert g.tobytes() == e.tobytes()


def test_prefix_chop_hand_derived_case(ctx):
    y = np.array([0, 10, 20], dtype=np.float32)
    x = np.array([0, 0, 10, 10, 20, 20, 100, 100, 100], dtype=np.float32)
    chop, hist, _ = ctx.prefix_chop(y, [x, np.concatenate([y, x])])
    assert int(chop[0]) == 4 and hist[0].tolist() == [2, 2, 1]


def test_prefix_chop_many_sequences_in_batches(ctx):
    """Enough sequences that the step words need several batches under a
    small memory cap."""
    rng = np.random.default_rng(44)
    prefix, seqs = _set(rng, 120, 300)
    ctx.set_mem_cap(64 << 20)
    try:
        chop, hist, _ = ctx.prefix_chop(prefix, seqs, norm=1)
    finally:
        ctx.set_mem_cap(1 << 62)
    echop, ehist, _ = _oracle(prefix, seqs, 1)
    assert np.array_equal(chop, echop)
    assert np.array_equal(hist, ehist)


def test_cli_prefix_file_end_to_end(tmp_path):
    """openDBA with a prefix file: <prefix>.prefix_chop.txt holds, per
    sequence, the name, chop position, original length and histogram of the
    oracle on z-normalised data, and avg.txt is the oracle's consensus of the
    chopped (already z-normalised) sequences."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "opendba_amd", "lib", "openDBA")
    assert os.path.exists(cli), "openDBA is not built"
    rng = np.random.default_rng(50)
    prefix, seqs = _set(rng, 5, 40)
    seqs = [s for s in seqs if len(s) > 60][:4] + [np.cumsum(rng.normal(size=700)).astype(np.float32)]
    files = []
    for k, s in enumerate(seqs):
        f = str(tmp_path / f"s{k}.txt")
        np.savetxt(f, s.astype(np.float64), fmt="%.9g")
        files.append(f)
    pf = str(tmp_path / "leader.txt")
    np.savetxt(pf, prefix.astype(np.float64), fmt="%.9g")
    out = str(tmp_path / "chop")
    print("cli: openDBA text float global with a prefix file", flush=True)
    r = subprocess.run([cli, "text", "float", "global", out, "0", pf, "1.0"] + files, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    read = [O.read_text_series(f, np.float32) for f in files]
    pre = O.read_text_series(pf, np.float32)
    echop, ehist, zseqs = _oracle(pre, read, 1)
    lines = open(out + ".prefix_chop.txt").read().rstrip("\n").split("\n")
    assert len(lines) == len(files)
    for k, ln in enumerate(lines):
        f = ln.split("\t")
        assert f[0] == files[k][: files[k].index(".")]
        assert int(f[1]) == int(echop[k]) and int(f[2]) == len(read[k])
        assert [int(v) for v in f[3:]] == ehist[k].tolist()
    chopped = [z[int(c):] for z, c in zip(zseqs, echop)]
    exp = O.perform_dba(chopped, norm=1)["centroid"]
    avg = open(out + ".avg.txt").read().strip().split("\n")[-1].split("\t")
    assert ["%.6g" % float(v) for v in avg[1:]] == ["%.6g" % v for v in exp]
