"""Writes tests/golden/slow5/expected.json from the reference's own BLOW5
test file (tests/good_files/slow5/FAN41461_pass_496845aa_0.blow5, copied
here as a fixture), decoded by this independent numpy reader of the file
format (record zlib, signal svb-zd: uint32 sample count, 2-bit control
codes for 1-4 little-endian bytes, zigzag-coded differences).

The decoding is pinned against the reference's edited-signal fixture
(tests/good_files/tsv/openDBA_test_edit_slow5.tsv): its three reads have
the decoded reads' ids and lengths, and their samples equal the decoded
samples at the same positions far more often than one position off (the
file holds edited copies of these signals), which a wrong decoding would
not give. Run in this container only (it reads /root/reference)."""
import json
import os
import struct
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests/good_files"
BLOW5 = os.path.join(HERE, "slow5", "FAN41461_pass_496845aa_0.blow5")


def fnv(sig):
    h = 1469598103934665603
    for b in sig.astype("<i2").tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def decode(path):
    d = open(path, "rb").read()
    assert d[:6] == b"BLOW5\x01" and d[9] == 1 and d[14] == 1  # zlib records, svb-zd signals
    hs = struct.unpack("<I", d[64:68])[0]
    p = 68 + hs
    out = []
    while d[p:p + 5] != b"5WOLB":
        sz = struct.unpack("<Q", d[p:p + 8])[0]
        r = zlib.decompress(d[p + 8:p + 8 + sz])
        p += 8 + sz
        ln = struct.unpack("<H", r[:2])[0]
        rid = r[2:2 + ln].decode()
        q = 2 + ln + 4 + 32
        nbytes = struct.unpack("<Q", r[q:q + 8])[0]
        n = struct.unpack("<I", r[q + 8:q + 12])[0]
        ncb = (n + 3) // 4
        ctrl = np.frombuffer(r[q + 12:q + 12 + ncb], dtype=np.uint8)
        codes = ((ctrl[:, None] >> (np.arange(4) * 2)) & 3).reshape(-1)[:n] + 1
        data = r[q + 12 + ncb:q + 8 + nbytes]
        assert int(codes.sum()) == len(data)
        ends = np.cumsum(codes)
        u = np.zeros(n, dtype=np.int64)
        for k in range(4):  # byte k of every value that has one
            has = codes > k
            u[has] |= np.frombuffer(data, dtype=np.uint8)[(ends - codes + k)[has]].astype(np.int64) << (8 * k)
        z = (u >> 1) ^ -(u & 1)
        out.append((rid, np.cumsum(z).astype(np.int16)))
    return out


def main():
    reads = decode(BLOW5)
    tsv = {}
    for line in open(os.path.join(REF, "tsv", "openDBA_test_edit_slow5.tsv")):
        f = line.rstrip("\n").split("\t")
        tsv[f[0]] = np.array([int(x) for x in f[1:] if x != ""])
    by = dict(reads)
    for rid, t in tsv.items():
        s = by[rid].astype(np.int64)
        assert len(s) == len(t), rid
        same = np.mean(s == t)
        off = max(np.mean(s[1:] == t[:-1]), np.mean(s[:-1] == t[1:]))
        print(f"{rid}: {len(s)} samples, equal at the same position {same:.3f}, one off {off:.3f}")
        assert same > 0.25 and same > 10 * off
    exp = [dict(id=rid, n=int(len(s)), fnv=fnv(s), head=[int(v) for v in s[:8]]) for rid, s in reads]
    json.dump(dict(file="FAN41461_pass_496845aa_0.blow5", reads=exp,
                   tsv_reads={k: int(len(v)) for k, v in tsv.items()}),
              open(os.path.join(HERE, "slow5", "expected.json"), "w"), indent=1)
    print(f"wrote {len(exp)} reads")


if __name__ == "__main__":
    main()
