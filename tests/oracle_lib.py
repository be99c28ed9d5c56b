"""ctypes wrapper around the CPU oracle (oracle/_build/libodba_oracle.so).

Test infrastructure: only tests, the smoke entry point and the benchmark's
CPU-baseline leg use it, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libodba_oracle.so")

I32, U32, U64, F32, F64 = 0, 1, 2, 3, 4
NP = {I32: np.int32, U32: np.uint32, U64: np.uint64, F32: np.float32, F64: np.float64}
CODE = {np.dtype(v): k for k, v in NP.items()}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(ORACLE_SO)
        vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
        L.orc_dtw_pair.argtypes = [i32, vp, u64, vp, u64, u64, i32, i32, vp, vp, vp, vp]
        L.orc_all_pairs.argtypes = [i32, vp, vp, u64, i32, i32, u64, vp, i32, vp, u64]
        L.orc_traceback_full.argtypes = [i32, vp, u64, u64, i32, vp, vp, vp, vp]
        L.orc_dba_update.argtypes = [i32, vp, u64, vp, vp, u64, i32, i32, i32, u64, vp, vp, vp]
        L.orc_znorm.argtypes = [i32, vp, vp, u64, vp, vp]
        L.orc_medoid_single.argtypes = [i32, vp, vp, u64, vp]
        L.orc_perform_dba.argtypes = [i32, vp, vp, u64, i32, i32, i32, i32, u64, i32, u64, vp, vp, vp, vp, vp, vp]
        L.orc_prefix_chop.argtypes = [i32, vp, u64, vp, u64, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ptrs(seqs):
    arr = (C.c_void_p * len(seqs))(*[s.ctypes.data for s in seqs])
    return arr


def dtype_code(a):
    return CODE[np.dtype(a.dtype)]


def dtw_pair(y, x, os_=0, oe=0, w=1024, moves=False):
    """Swath-faithful DTW of one pair; returns (distance, moves or None, abrogated swath, q3 flag)."""
    y = np.ascontiguousarray(y)
    x = np.ascontiguousarray(x)
    assert y.dtype == x.dtype
    dt = dtype_code(y)
    d = np.zeros(1, dtype=y.dtype)
    M = np.zeros((len(y), len(x)), dtype=np.uint8) if moves else None
    ab = C.c_int(0)
    q3 = C.c_int(0)
    rc = lib().orc_dtw_pair(dt, _p(y), len(y), _p(x), len(x), w, os_, oe, _p(M) if moves else None, _p(d),
                            C.byref(ab), C.byref(q3))
    assert rc == 0, rc
    return d[0], M, ab.value, q3.value


def all_pairs(seqs, os_=0, oe=0, w=1024, nthreads=8, subset=None):
    seqs = [np.ascontiguousarray(s) for s in seqs]
    dt = dtype_code(seqs[0])
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    if subset is None:
        out = np.zeros(n * (n - 1) // 2, dtype=seqs[0].dtype)
        sub, nsub = None, 0
    else:
        subset = np.ascontiguousarray(subset, dtype=np.int64)
        out = np.zeros(len(subset), dtype=seqs[0].dtype)
        sub, nsub = _p(subset), len(subset)
    rc = lib().orc_all_pairs(dt, _ptrs(seqs), _p(lens), n, os_, oe, w, _p(out), nthreads, sub, nsub)
    assert rc == 0, rc
    return out


def prefix_chop(prefix, seq):
    """chopPrefixFromSequences for one sequence (inputs already normalised):
    returns (chop position, per-prefix-row path-cell histogram)."""
    prefix = np.ascontiguousarray(prefix)
    seq = np.ascontiguousarray(seq)
    assert prefix.dtype == seq.dtype
    chop = C.c_uint64(0)
    hist = np.zeros(len(prefix), dtype=np.int32)
    rc = lib().orc_prefix_chop(dtype_code(seq), _p(prefix), len(prefix), _p(seq), len(seq), C.byref(chop), _p(hist))
    assert rc == 0, rc
    return chop.value, hist


def traceback_full(moves, flip, seq, lc):
    seq = np.ascontiguousarray(seq)
    sums = np.zeros(lc, dtype=seq.dtype)
    counts = np.zeros(lc, dtype=np.uint32)
    pl = C.c_uint64(0)
    rows, cols = moves.shape
    rc = lib().orc_traceback_full(dtype_code(seq), _p(np.ascontiguousarray(moves)), rows, cols, flip, _p(seq),
                                  _p(sums), _p(counts), C.byref(pl))
    return rc, sums, counts, pl.value


def dba_update(centroid, seqs, os_=0, oe=0, path_policy=1, free_mem=0):
    centroid = np.ascontiguousarray(centroid)
    seqs = [np.ascontiguousarray(s) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    out = np.zeros_like(centroid)
    counts = np.zeros(len(centroid), dtype=np.uint32)
    delta = C.c_double(0)
    rc = lib().orc_dba_update(dtype_code(centroid), _p(centroid), len(centroid), _ptrs(seqs), _p(lens), len(seqs),
                              os_, oe, path_policy, free_mem, _p(out), _p(counts), C.byref(delta))
    return rc, out, counts, delta.value


def znorm(seqs):
    seqs = [np.array(s, copy=True) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    mu = np.zeros(len(seqs))
    sg = np.zeros(len(seqs))
    rc = lib().orc_znorm(dtype_code(seqs[0]), _ptrs(seqs), _p(lens), len(seqs), _p(mu), _p(sg))
    assert rc == 0
    return seqs, mu, sg


def medoid_single(dists, lens):
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    sos = np.zeros(len(lens), dtype=dists.dtype)
    m = lib().orc_medoid_single(dtype_code(dists), _p(np.ascontiguousarray(dists)), _p(lens), len(lens), _p(sos))
    return m, sos


def perform_dba(seqs, os_=0, oe=0, norm=0, max_rounds=250, medoid_w=1024, path_policy=1, free_mem=0):
    seqs = [np.array(s, copy=True) for s in seqs]
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    cen = np.zeros(int(lens.max()), dtype=seqs[0].dtype)
    clen = C.c_uint64(0)
    med = C.c_int(0)
    perm = np.zeros(n, dtype=np.int32)
    rounds = C.c_int(0)
    dists = np.zeros(n * (n - 1) // 2, dtype=seqs[0].dtype)
    rc = lib().orc_perform_dba(dtype_code(seqs[0]), _ptrs(seqs), _p(lens), n, os_, oe, norm, max_rounds, medoid_w,
                               path_policy, free_mem, _p(cen), C.byref(clen), C.byref(med), _p(perm),
                               C.byref(rounds), _p(dists))
    assert rc == 0, rc
    return dict(centroid=cen[:clen.value], medoid=med.value, perm=perm, rounds=rounds.value, dists=dists)


_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.c_void_p]


def read_text_series(path, dtype=np.float32):
    """readSequenceTextFiles semantics (cpu_utils.hpp:92-134): one value per
    line, the value count is the number of newlines, parsed with the C
    library's correctly rounded strtof / strtod."""
    with open(path, "rb") as f:
        data = f.read()
    n = data.count(b"\n")
    lines = data.split(b"\n")[:n]
    out = np.zeros(n, dtype=dtype)
    for k, ln in enumerate(lines):
        tok = ln.split()[0] if ln.split() else b"0"
        if dtype == np.float32:
            out[k] = _libc.strtof(tok, None)
        elif dtype == np.float64:
            out[k] = _libc.strtod(tok, None)
        else:
            out[k] = int(float(tok)) if b"." not in tok else int(tok.split(b".")[0] or b"0")
    return out


def read_tokens(path, dtype=np.float64):
    """All whitespace-separated numbers of a file (for expected-output fixtures,
    which are not read through the reference's reader)."""
    with open(path, "rb") as f:
        toks = f.read().split()
    conv = _libc.strtof if dtype == np.float32 else _libc.strtod
    return np.array([conv(t, None) for t in toks], dtype=dtype)


def kseg_params(minseg):
    """ds, samples per block and max segments per block (segmentation.hpp:347-353)."""
    ds = (minseg + 2) // 3
    spb = 256 * ds
    x = np.float32(minseg) / np.float32(ds)
    mk = int((np.float32(256) + x - np.float32(1)) / x)
    return ds, spb, mk


def kseg(seqs, minseg, prefix_skip=0, lds_bytes=65536):
    seqs = [np.ascontiguousarray(s) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    ds, spb, mk = kseg_params(minseg)
    cap = int(sum(mk * ((int(L) + spb - 1) // spb) for L in lens))
    out = np.zeros(max(cap, 1), dtype=seqs[0].dtype)
    ol = np.zeros(len(seqs), dtype=np.uint64)
    oo = np.zeros(len(seqs), dtype=np.uint64)
    L = lib()
    L.orc_kseg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_void_p,
                           C.c_void_p, C.c_void_p]
    rc = L.orc_kseg(dtype_code(seqs[0]), _ptrs(seqs), _p(lens), len(seqs), minseg, prefix_skip, lds_bytes, _p(out),
                    _p(ol), _p(oo))
    assert rc == 0, rc
    return [out[int(oo[i]):int(oo[i]) + int(ol[i])].copy() for i in range(len(seqs))]
