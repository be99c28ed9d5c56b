"""S/BLOW5 signal files of the drop-in layer (include/odba/odba_slow5.hpp;
the reference reads and writes them through slow5lib: cpu_utils.hpp:136-278,
458-500, io_utils.hpp:363-452). Checked on the reference's own BLOW5 test
file against tests/golden/slow5/expected.json, whose independent decoding is
pinned to the reference's edited-signal fixture (make_slow5_fixture.py);
CPU only."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "slow5")
BLOW5 = os.path.join(GOLD, "FAN41461_pass_496845aa_0.blow5")
TOOL = os.path.join(ROOT, "tests", "cpp", "_build", "slow5_tool")


@pytest.fixture(scope="module")
def tool():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "_build/slow5_tool"])
    return TOOL


def dump(tool, path):
    out = subprocess.run([tool, "dump", path], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    rows = []
    for line in out.stdout.splitlines():
        rid, n, h, head = line.split("\t")
        rows.append(dict(id=rid, n=int(n), fnv=h, head=[int(v) for v in head.split(",")]))
    return rows


def test_blow5_reads_match_the_pinned_decoding(tool):
    exp = json.load(open(os.path.join(GOLD, "expected.json")))
    got = dump(tool, BLOW5)
    assert got == exp["reads"]
    names = {r["id"]: r["n"] for r in got}
    for rid, n in exp["tsv_reads"].items():  # the reference's edit fixture names these reads
        assert names[rid] == n


def test_blow5_rewrite_round_trip(tool, tmp_path):
    """writeSlow5Output with new signals of the same lengths, in reverse read
    order: the copy keeps the header and every other field of each record
    byte for byte, holds the new signals, and re-encodes in the input's
    compression (zlib records, svb-zd signals)."""
    out = str(tmp_path / "new.blow5")
    r = subprocess.run([tool, "rewrite", BLOW5, out, "7"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok 11", (r.returncode, r.stdout, r.stderr)
    rows = dump(tool, out)
    assert [x["id"] for x in rows] == [x["id"] for x in reversed(dump(tool, BLOW5))]


@pytest.mark.parametrize("name,length", [("no-such-read", 10), ("582addd8-7822-4236-82e0-32dc6559f276", 42346)])
def test_blow5_rewrite_refuses_unknown_names_and_lengths(tool, tmp_path, name, length):
    """As the reference does (io_utils.hpp:405-421): a name missing from the
    file, or a length different from the read's, fails with return 1."""
    r = subprocess.run([tool, "rewrite_named", BLOW5, str(tmp_path / "x.blow5"), name, str(length)],
                       capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "1", (r.stdout, r.stderr)


def test_slow5_text_records(tool, tmp_path):
    """The text form: header lines, column lines, then tab-separated records
    with the signal comma-separated."""
    rng = np.random.default_rng(3)
    sigs = [rng.integers(-500, 1500, size=n).astype(np.int16) for n in (5, 1, 300)]
    p = tmp_path / "t.slow5"
    with open(p, "w") as f:
        f.write("#slow5_version\t0.2.0\n#num_read_groups\t1\n@run_id\tx\n")
        f.write("#char*\tuint32_t\tdouble\tdouble\tdouble\tdouble\tuint64_t\tint16_t*\tuint8_t\n")
        f.write("#read_id\tread_group\tdigitisation\toffset\trange\tsampling_rate\tlen_raw_signal\traw_signal\tstart_mux\n")
        for k, s in enumerate(sigs):
            f.write(f"read{k}\t0\t8192\t12\t1106.6\t3012\t{len(s)}\t{','.join(str(int(v)) for v in s)}\t1\n")
    got = dump(tool, str(p))
    assert [g["id"] for g in got] == ["read0", "read1", "read2"]
    for g, s in zip(got, sigs):
        assert g["n"] == len(s) and g["head"] == [int(v) for v in s[:8]]


def test_corrupt_blow5_is_unreadable(tool, tmp_path):
    data = bytearray(open(BLOW5, "rb").read())
    data[5000:5100] = bytes(100)  # inside the first record's compressed bytes
    p = tmp_path / "bad.blow5"
    p.write_bytes(bytes(data))
    r = subprocess.run([tool, "dump", str(p)], capture_output=True, text=True)
    assert r.returncode == 41  # SLOW5_FILE_UNREADABLE (exit_codes.hpp:27)
