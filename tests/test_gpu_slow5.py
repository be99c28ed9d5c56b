"""The openDBA command line on the reference's own BLOW5 test file
(tests/good_files/slow5/FAN41461_pass_496845aa_0.blow5, a fixture here):
'slow5' read mode (openDBA.cu:54-57, cpu_utils.hpp:458-500), and the
consensus written back into a copy of the medoid's record as
<prefix>.avg.blow5 (dba.hpp:1170-1185, io_utils.hpp:363-452). The signals
the CLI reads are those of the pinned decoding (tests/golden/slow5); the
consensus equals the library's perform_dba on them, and the BLOW5 record
holds its short conversion."""
import os
import subprocess
import sys

import numpy as np
import pytest

from opendba_amd import native as N

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BLOW5 = os.path.join(GOLD, "slow5", "FAN41461_pass_496845aa_0.blow5")
CLI = os.path.join(ROOT, "opendba_amd", "lib", "openDBA")
TOOL = os.path.join(ROOT, "tests", "cpp", "_build", "slow5_tool")
sys.path.insert(0, GOLD)
import make_slow5_fixture as F  # noqa: E402  (the independent decoder)


def fnv(a):
    return F.fnv(np.asarray(a, dtype=np.int16))


def test_cli_slow5_consensus_and_avg_blow5(tmp_path):
    reads = F.decode(BLOW5)
    prefix = str(tmp_path / "out")
    print("run: openDBA slow5 float open_end", flush=True)
    r = subprocess.run([CLI, "slow5", "float", "open_end", prefix, "0", "/dev/null", "1", BLOW5],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = open(prefix + ".avg.txt").read().rstrip("\n").split("\n")
    assert len(lines) == 1
    name, *vals = lines[0].split("\t")
    got = np.array([float(v) for v in vals])
    # the same consensus through the library (same reads, same order)
    seqs = [s.astype(np.float32) for _, s in reads]
    with N.Context(0) as ctx:
        res = ctx.perform_dba(seqs, 0, 1, norm=1, max_rounds=250)
    names = [rid for rid, _ in reads]
    med_name = names[int(res["perm"][res["medoid"]])]
    assert name == med_name
    cen = res["centroid"]
    assert len(got) == len(cen)
    np.testing.assert_allclose(got, cen, rtol=1e-5, atol=1e-3)  # avg.txt prints 6 significant digits
    # the BLOW5 copy: the medoid's record with the consensus as shorts
    d = subprocess.run([TOOL, "dump", prefix + ".avg.blow5"], capture_output=True, text=True)
    assert d.returncode == 0, d.stderr
    rows = [ln.split("\t") for ln in d.stdout.splitlines()]
    assert len(rows) == 1 and rows[0][0] == med_name and int(rows[0][1]) == len(cen)
    assert rows[0][2] == fnv(cen.astype(np.int16))  # (short) of each float, truncating as the reference's cast
