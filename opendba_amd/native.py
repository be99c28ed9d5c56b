"""ctypes binding over libodba_hip.so (include/odba/odba.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libodba_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "odba", "odba.h")

I32, U32, U64, F32, F64 = 0, 1, 2, 3, 4
NP = {I32: np.int32, U32: np.uint32, U64: np.uint64, F32: np.float32, F64: np.float64}
CODE = {np.dtype(v): k for k, v in NP.items()}
PATH_AUTO, PATH_FULL, PATH_STRIPE = 0, 1, 2

EXPORTS = [
    "odba_ctx_create", "odba_ctx_destroy", "odba_last_error", "odba_device_info",
    "odba_dtw_all_pairs", "odba_seqset_create", "odba_seqset_destroy", "odba_dtw_all_pairs_resident",
    "odba_dtw_path", "odba_dba_round", "odba_znorm", "odba_perform_dba", "odba_kseg", "odba_kseg_capacity",
    "odba_dba_create", "odba_dba_destroy", "odba_dba_set_centroid", "odba_dba_step",
    "odba_dba_record_paths", "odba_dba_get_path", "odba_ctx_set_mem_cap",
    "odba_server_create", "odba_server_destroy", "odba_server_submit", "odba_server_wait", "odba_server_result",
    "odba_server_release", "odba_server_stats", "odba_dba_clusters", "odba_cluster", "odba_hclust_complete",
    "odba_cutree_k", "odba_prefix_chop", "odba_tune_all_pairs", "odba_get_schedule", "odba_set_schedule",
    "odba_ctx_set_graphs", "odba_dtw_all_pairs_noalloc", "odba_job_prepare", "odba_job_result", "odba_job_destroy",
    "odba_job_words", "odba_job_shape",
    "odba_rounds_create", "odba_rounds_add", "odba_rounds_step", "odba_rounds_destroy",
]
SCHED_WAVE, SCHED_BANDS = 0, 1


class Schedule(C.Structure):
    """odba_schedule (odba_types.h)."""
    _fields_ = [("kind", C.c_int), ("order", C.c_int), ("band_c", C.c_int), ("band_waves", C.c_int),
                ("ring", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


class OdbaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"odba error {code}: {msg}")
        self.code = code


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OdbaError(-1, f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C opendba_amd/csrc")
    L = C.CDLL(LIB_PATH)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    L.odba_ctx_create.argtypes = [i32, C.POINTER(vp)]
    L.odba_ctx_destroy.argtypes = [vp]
    L.odba_last_error.argtypes = [vp]
    L.odba_last_error.restype = C.c_char_p
    L.odba_device_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(u64), C.c_char_p, i32]
    L.odba_dtw_all_pairs.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, vp]
    L.odba_seqset_create.argtypes = [vp, i32, vp, vp, i32, C.POINTER(vp)]
    L.odba_seqset_destroy.argtypes = [vp]
    L.odba_dtw_all_pairs_resident.argtypes = [vp, vp, i32, i32, vp, vp, C.POINTER(C.c_double)]
    L.odba_dtw_all_pairs_noalloc.argtypes = [vp, vp, i32, i32, vp, vp]
    L.odba_dtw_path.argtypes = [vp, i32, vp, u64, vp, u64, i32, i32, vp, vp]
    L.odba_dba_round.argtypes = [vp, i32, vp, u64, vp, vp, i32, i32, i32, i32, u64, vp, vp, C.POINTER(C.c_double)]
    L.odba_znorm.argtypes = [vp, i32, vp, vp, i32, vp, vp]
    L.odba_perform_dba.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, i32, u64, vp, C.POINTER(u64),
                                   C.POINTER(i32), vp, C.POINTER(i32), vp]
    L.odba_kseg.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, vp, u64, vp, vp]
    L.odba_kseg_capacity.argtypes = [vp, i32, i32]
    L.odba_kseg_capacity.restype = u64
    L.odba_dba_create.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, u64, C.POINTER(vp)]
    L.odba_dba_destroy.argtypes = [vp]
    L.odba_dba_set_centroid.argtypes = [vp, vp, u64]
    L.odba_dba_step.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
    L.odba_dba_record_paths.argtypes = [vp, i32]
    L.odba_dba_get_path.argtypes = [vp, i32, vp, u64, C.POINTER(u64)]
    L.odba_ctx_set_mem_cap.argtypes = [vp, u64]
    L.odba_ctx_set_graphs.argtypes = [vp, i32]
    L.odba_prefix_chop.argtypes = [vp, i32, vp, u64, vp, vp, i32, i32, vp, vp]
    L.odba_tune_all_pairs.argtypes = [vp, vp, i32, i32, i32, C.POINTER(Schedule), C.POINTER(C.c_double),
                                      C.POINTER(i32)]
    L.odba_get_schedule.argtypes = [vp, vp, i32, i32, C.POINTER(Schedule)]
    L.odba_set_schedule.argtypes = [vp, vp, i32, i32, C.POINTER(Schedule)]
    L.odba_server_create.argtypes = [i32, u64, i32, C.POINTER(vp)]
    L.odba_server_destroy.argtypes = [vp]
    L.odba_server_submit.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, C.POINTER(u64)]
    L.odba_server_wait.argtypes = [vp, u64, C.c_double, C.POINTER(i32)]
    L.odba_server_result.argtypes = [vp, u64, vp, u64, C.POINTER(u64), C.POINTER(i32), C.POINTER(i32),
                                     C.POINTER(C.c_double), C.POINTER(i32)]
    L.odba_server_release.argtypes = [vp, u64]
    L.odba_server_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32), C.POINTER(i32)]
    L.odba_dba_clusters.argtypes = [vp, i32, vp, vp, i32, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp, vp]
    L.odba_cluster.argtypes = [i32, vp, i32, vp, C.c_double, vp, vp, C.POINTER(i32)]
    L.odba_hclust_complete.argtypes = [i32, vp, vp, vp]
    L.odba_cutree_k.argtypes = [i32, vp, i32, vp]
    L.odba_job_prepare.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, C.POINTER(vp)]
    L.odba_job_result.argtypes = [vp, vp, u64, C.POINTER(u64), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.odba_job_destroy.argtypes = [vp]
    L.odba_job_words.argtypes = [vp, i32, C.POINTER(u64)]
    L.odba_job_shape.argtypes = [vp, C.POINTER(u64), C.POINTER(i32)]
    L.odba_rounds_create.argtypes = [vp, i32, i32, i32, C.POINTER(vp)]
    L.odba_rounds_add.argtypes = [vp, vp]
    L.odba_rounds_step.argtypes = [vp, vp, i32, C.POINTER(i32), C.POINTER(i32)]
    L.odba_rounds_destroy.argtypes = [vp]
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ptrs(seqs):
    return (C.c_void_p * len(seqs))(*[s.ctypes.data for s in seqs])


class Context:
    """One HIP device + stream (odba_ctx)."""

    def __init__(self, device=0):
        self.L = load()
        self.h = C.c_void_p()
        rc = self.L.odba_ctx_create(device, C.byref(self.h))
        if rc:
            raise OdbaError(rc, "odba_ctx_create failed (no HIP device?)")

    def close(self):
        if self.h:
            self.L.odba_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            raise OdbaError(rc, self.L.odba_last_error(self.h).decode())

    def device_info(self):
        cu, clk, mem = C.c_int(), C.c_int(), C.c_uint64()
        name = C.create_string_buffer(64)
        self._check(self.L.odba_device_info(self.h, C.byref(cu), C.byref(clk), C.byref(mem), name, 64))
        return dict(compute_units=cu.value, clock_khz=clk.value, mem_bytes=mem.value, name=name.value.decode())

    def all_pairs(self, seqs, open_start=0, open_end=0):
        seqs = [np.ascontiguousarray(s) for s in seqs]
        dt = CODE[seqs[0].dtype]
        n = len(seqs)
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        out = np.zeros(n * (n - 1) // 2, dtype=seqs[0].dtype)
        self._check(self.L.odba_dtw_all_pairs(self.h, dt, _ptrs(seqs), _p(lens), n, open_start, open_end, 1024,
                                              _p(out)))
        return out

    def seqset(self, seqs):
        return SeqSet(self, seqs)

    def dtw_path(self, y, x, open_start=0, open_end=0):
        y = np.ascontiguousarray(y)
        x = np.ascontiguousarray(x)
        M = np.zeros((len(y), len(x)), dtype=np.uint8)
        d = np.zeros(1, dtype=y.dtype)
        self._check(self.L.odba_dtw_path(self.h, CODE[y.dtype], _p(y), len(y), _p(x), len(x), open_start, open_end,
                                         _p(M), _p(d)))
        return d[0], M

    def dba_round(self, centroid, seqs, open_start=0, open_end=0, path_policy=PATH_FULL, ref_free_mem=0):
        centroid = np.ascontiguousarray(centroid)
        seqs = [np.ascontiguousarray(s) for s in seqs]
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        out = np.zeros_like(centroid)
        counts = np.zeros(len(centroid), dtype=np.uint32)
        delta = C.c_double(0)
        self._check(self.L.odba_dba_round(self.h, CODE[centroid.dtype], _p(centroid), len(centroid), _ptrs(seqs),
                                          _p(lens), len(seqs), open_start, open_end, path_policy, ref_free_mem,
                                          _p(out), _p(counts), C.byref(delta)))
        return out, counts, delta.value

    def znorm(self, seqs):
        seqs = [np.array(s, copy=True) for s in seqs]
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        mu = np.zeros(len(seqs))
        sg = np.zeros(len(seqs))
        self._check(self.L.odba_znorm(self.h, CODE[seqs[0].dtype], _ptrs(seqs), _p(lens), len(seqs), _p(mu),
                                      _p(sg)))
        return seqs, mu, sg

    def set_mem_cap(self, nbytes):
        """Device memory the context plans its batches within (odba_ctx_set_mem_cap)."""
        self._check(self.L.odba_ctx_set_mem_cap(self.h, int(nbytes)))

    def prefix_chop(self, prefix, seqs, norm=0):
        """Prefix location (odba_prefix_chop): returns (chop positions,
        histograms [n, len(prefix)], the sequences as the library left them
        (z-normalised copies when norm))."""
        prefix = np.ascontiguousarray(prefix)
        seqs = [np.array(s, copy=True) for s in seqs]
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        chop = np.zeros(len(seqs), dtype=np.uint64)
        hist = np.zeros((len(seqs), len(prefix)), dtype=np.int32)
        self._check(self.L.odba_prefix_chop(self.h, CODE[prefix.dtype], _p(prefix), len(prefix), _ptrs(seqs), _p(lens),
                                            len(seqs), norm, _p(chop), _p(hist)))
        return chop, hist, seqs

    def dba_clusters(self, seqs, memberships, medoids, open_start=0, open_end=0, max_rounds=250, mu=None,
                     sigma=None):
        """All clusters' DBA with packed rounds (odba_dba_clusters); returns
        (list of centroids, rounds per cluster)."""
        seqs = [np.ascontiguousarray(s) for s in seqs]
        n = len(seqs)
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        memb = np.ascontiguousarray(memberships, dtype=np.int32)
        med = np.ascontiguousarray(medoids, dtype=np.int32)
        k = len(med)
        offs = np.zeros(k + 1, dtype=np.uint64)
        out = np.zeros(int(sum(len(seqs[m]) for m in med)), dtype=seqs[0].dtype)
        rounds = np.zeros(k, dtype=np.int32)
        mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        sigma = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        self._check(self.L.odba_dba_clusters(self.h, CODE[seqs[0].dtype], _ptrs(seqs), _p(lens), n, _p(memb), _p(med),
                                             k, open_start, open_end, max_rounds, _p(mu), _p(sigma), _p(out),
                                             _p(offs), _p(rounds)))
        return [out[int(offs[i]):int(offs[i + 1])] for i in range(k)], rounds

    def perform_dba(self, seqs, open_start=0, open_end=0, norm=0, max_rounds=250, path_policy=PATH_FULL,
                    ref_free_mem=0):
        seqs = [np.ascontiguousarray(s) for s in seqs]
        n = len(seqs)
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        cen = np.zeros(int(lens.max()), dtype=seqs[0].dtype)
        clen = C.c_uint64()
        med = C.c_int()
        perm = np.zeros(n, dtype=np.int32)
        rounds = C.c_int()
        dists = np.zeros(n * (n - 1) // 2, dtype=seqs[0].dtype)
        self._check(self.L.odba_perform_dba(self.h, CODE[seqs[0].dtype], _ptrs(seqs), _p(lens), n, open_start,
                                            open_end, norm, max_rounds, path_policy, ref_free_mem, _p(cen),
                                            C.byref(clen), C.byref(med), _p(perm), C.byref(rounds), _p(dists)))
        return dict(centroid=cen[: clen.value], medoid=med.value, perm=perm, rounds=rounds.value, dists=dists)


def cluster(dists, lens, cdist):
    """Clusters and medoids from all-vs-all distances (odba_cluster; host
    only). Returns (memberships, medoids)."""
    L = load()
    dists = np.ascontiguousarray(dists)
    n = len(lens)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    memb = np.zeros(n, dtype=np.int32)
    med = np.zeros(max(n, 1), dtype=np.int32)
    k = C.c_int()
    rc = L.odba_cluster(CODE[dists.dtype], _p(dists), n, _p(lens), float(cdist), _p(memb), _p(med), C.byref(k))
    if rc:
        raise OdbaError(rc, "odba_cluster failed")
    return memb, med[: k.value]


def hclust_complete(condensed, n):
    L = load()
    d = np.ascontiguousarray(condensed, dtype=np.float64)
    merge = np.zeros(2 * (n - 1), dtype=np.int32)
    height = np.zeros(n - 1, dtype=np.float64)
    rc = L.odba_hclust_complete(n, _p(d), _p(merge), _p(height))
    if rc:
        raise OdbaError(rc, "odba_hclust_complete failed")
    return merge.reshape(2, n - 1).T.copy(), height


def cutree_k(merge, nclust):
    L = load()
    n = merge.shape[0] + 1
    m = np.ascontiguousarray(merge.T.reshape(-1), dtype=np.int32)
    labels = np.zeros(n, dtype=np.int32)
    rc = L.odba_cutree_k(n, _p(m), nclust, _p(labels))
    if rc:
        raise OdbaError(rc, "odba_cutree_k failed")
    return labels


def unpack_steps(packed):
    """(i, j, code) arrays of packed traceback steps; code is the reference
    step code (NIL = 255)."""
    packed = np.asarray(packed, dtype=np.uint64)
    i = (packed >> np.uint64(33)).astype(np.int64)
    j = ((packed >> np.uint64(3)) & np.uint64(0x3FFFFFFF)).astype(np.int64)
    code = (packed & np.uint64(7)).astype(np.uint8)
    code[code == 0] = 255
    return i, j, code


class DbaSession:
    """Device-resident DBA session (odba_dba_*): sequences uploaded once,
    one DBAUpdate per step."""

    def __init__(self, ctx, seqs, open_start=0, open_end=0, path_policy=PATH_FULL, ref_free_mem=0):
        self.ctx = ctx
        self.L = ctx.L
        self.seqs = [np.ascontiguousarray(s) for s in seqs]
        self.dtype = self.seqs[0].dtype
        lens = np.array([len(s) for s in self.seqs], dtype=np.uint64)
        self.h = C.c_void_p()
        ctx._check(self.L.odba_dba_create(ctx.h, CODE[self.dtype], _ptrs(self.seqs), _p(lens), len(self.seqs),
                                          open_start, open_end, path_policy, ref_free_mem, C.byref(self.h)))
        self.lc = 0

    def set_centroid(self, c):
        c = np.ascontiguousarray(c, dtype=self.dtype)
        self.lc = len(c)
        self.ctx._check(self.L.odba_dba_set_centroid(self.h, _p(c), len(c)))

    def record_paths(self, on=True):
        self.ctx._check(self.L.odba_dba_record_paths(self.h, 1 if on else 0))

    def step(self):
        out = np.zeros(self.lc, dtype=self.dtype)
        counts = np.zeros(self.lc, dtype=np.uint32)
        delta = C.c_double(0)
        self.ctx._check(self.L.odba_dba_step(self.h, _p(out), _p(counts), C.byref(delta)))
        return out, counts, delta.value

    def path(self, seq):
        n = C.c_uint64()
        self.ctx._check(self.L.odba_dba_get_path(self.h, seq, None, 0, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint64)
        self.ctx._check(self.L.odba_dba_get_path(self.h, seq, _p(buf), n.value, C.byref(n)))
        return unpack_steps(buf)

    def close(self):
        if self.h:
            self.L.odba_dba_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _kseg(ctx, seqs, min_segment_length, prefix_skip=0, lds_bytes=65536):
    seqs = [np.ascontiguousarray(s) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    cap = int(ctx.L.odba_kseg_capacity(_p(lens), len(seqs), min_segment_length))
    out = np.zeros(max(cap, 1), dtype=seqs[0].dtype)
    ol = np.zeros(len(seqs), dtype=np.uint64)
    oo = np.zeros(len(seqs), dtype=np.uint64)
    ctx._check(ctx.L.odba_kseg(ctx.h, CODE[seqs[0].dtype], _ptrs(seqs), _p(lens), len(seqs), min_segment_length,
                               prefix_skip, lds_bytes, _p(out), cap, _p(ol), _p(oo)))
    return [out[int(oo[i]):int(oo[i]) + int(ol[i])].copy() for i in range(len(seqs))]


Context.kseg = _kseg


class SeqSet:
    """Device-resident sequence set (odba_seqset)."""

    def __init__(self, ctx, seqs):
        self.ctx = ctx
        seqs = [np.ascontiguousarray(s) for s in seqs]
        self.dtype = seqs[0].dtype
        self.n = len(seqs)
        self.lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        self.h = C.c_void_p()
        ctx._check(ctx.L.odba_seqset_create(ctx.h, CODE[self.dtype], _ptrs(seqs), _p(self.lens), self.n,
                                            C.byref(self.h)))

    def all_pairs(self, open_start=0, open_end=0, copy_out=True, flags=False):
        n = self.n
        out = np.zeros(n * (n - 1) // 2, dtype=self.dtype) if copy_out else None
        fl = np.zeros(n * (n - 1) // 2, dtype=np.int32) if flags else None
        ms = C.c_double(0)
        self.ctx._check(self.ctx.L.odba_dtw_all_pairs_resident(self.ctx.h, self.h, open_start, open_end, _p(out),
                                                               _p(fl), C.byref(ms)))
        return out, fl, ms.value

    def tune(self, open_start=0, open_end=0, reps=3):
        """Autotune the all-vs-all schedule for this set and mode
        (odba_tune_all_pairs); returns (schedule dict, kernel ms, candidates)."""
        best = Schedule()
        ms = C.c_double(0)
        tried = C.c_int(0)
        self.ctx._check(self.ctx.L.odba_tune_all_pairs(self.ctx.h, self.h, open_start, open_end, reps, C.byref(best),
                                                       C.byref(ms), C.byref(tried)))
        return best.as_dict(), ms.value, tried.value

    def schedule(self, open_start=0, open_end=0):
        out = Schedule()
        self.ctx._check(self.ctx.L.odba_get_schedule(self.ctx.h, self.h, open_start, open_end, C.byref(out)))
        return out.as_dict()

    def set_schedule(self, sched, open_start=0, open_end=0):
        """Pin a schedule (a dict with odba_schedule's fields) or, with None,
        go back to the automatic choice."""
        if sched is None:
            self.ctx._check(self.ctx.L.odba_set_schedule(self.ctx.h, self.h, open_start, open_end, None))
            return
        ring = 0 if int(sched.get("kind", 0)) == 2 else 2048  # small: 0 = 16-step sub-blocks
        u = Schedule(**{k: int(sched.get(k, d)) for k, d in
                        (("kind", 0), ("order", 0), ("band_c", 16), ("band_waves", 4), ("ring", ring))})
        self.ctx._check(self.ctx.L.odba_set_schedule(self.ctx.h, self.h, open_start, open_end, C.byref(u)))

    def close(self):
        if self.h:
            self.ctx.L.odba_seqset_destroy(self.h)
            self.h = C.c_void_p()


class Job:
    """A prepared DBA job (odba_job_prepare): sanitized, sorted, z-normalised
    sequences resident on the device, all-vs-all done, medoid found."""

    def __init__(self, ctx, seqs, open_start=0, open_end=0, norm=1, max_rounds=250):
        self.L = ctx.L
        self.ctx = ctx
        seqs = [np.ascontiguousarray(s) for s in seqs]
        self.dtype = seqs[0].dtype
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        self.maxlen = int(lens.max())
        self.h = C.c_void_p()
        ctx._check(self.L.odba_job_prepare(ctx.h, CODE[self.dtype], _ptrs(seqs), _p(lens), len(seqs), open_start,
                                           open_end, norm, max_rounds, C.byref(self.h)))

    def result(self):
        cen = np.zeros(self.maxlen, dtype=self.dtype)
        clen = C.c_uint64()
        med, rounds, jrc = C.c_int(), C.c_int(), C.c_int()
        rc = self.L.odba_job_result(self.h, _p(cen), self.maxlen, C.byref(clen), C.byref(med), C.byref(rounds),
                                    C.byref(jrc))
        if rc:
            raise OdbaError(rc, "odba_job_result failed")
        return dict(centroid=cen[: clen.value], medoid=med.value, rounds=rounds.value, rc=jrc.value)

    def words(self, allocate=False):
        """Bytes of the job's step words (odba_job_words); with allocate, the
        buffer is allocated now."""
        wb = C.c_uint64()
        rc = self.L.odba_job_words(self.h, 1 if allocate else 0, C.byref(wb))
        if rc:
            raise OdbaError(rc, "odba_job_words failed")
        return wb.value

    def shape(self):
        """(centroid length, largest swath count of its path matrices)."""
        lc, k = C.c_uint64(), C.c_int()
        rc = self.L.odba_job_shape(self.h, C.byref(lc), C.byref(k))
        if rc:
            raise OdbaError(rc, "odba_job_shape failed")
        return lc.value, k.value

    def close(self):
        if self.h:
            self.L.odba_job_destroy(self.h)
            self.h = C.c_void_p()


class Rounds:
    """Packed DBA rounds over prepared jobs of one value type and mode
    (odba_rounds_*): each step runs one round of every job held."""

    def __init__(self, ctx, dtype, open_start=0, open_end=0):
        self.L = ctx.L
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx._check(self.L.odba_rounds_create(ctx.h, CODE[np.dtype(dtype)], open_start, open_end, C.byref(self.h)))

    def add(self, job):
        self.ctx._check(self.L.odba_rounds_add(self.h, job.h))

    def step(self, capacity=256):
        """One packed round; returns (handles of jobs that ended, jobs still held)."""
        fin = (C.c_void_p * capacity)()
        nf, na = C.c_int(), C.c_int()
        self.ctx._check(self.L.odba_rounds_step(self.h, fin, capacity, C.byref(nf), C.byref(na)))
        return [fin[k] for k in range(nf.value)], na.value

    def close(self):
        if self.h:
            self.L.odba_rounds_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
