"""Serving DBA jobs on one GPU, and the plain per-job path it is measured
against.

`Server` wraps the library's serving loop (odba_server_*): jobs are admitted
in arrival order when a worker slot is free and their estimated device
working set fits in the HBM budget, and run concurrently on separate HIP
streams. `run_trace` replays a seeded Poisson trace against either the
server or the plain path (one odba_perform_dba after another, in arrival
order) and reports throughput and latency percentiles.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import native as N


class Server:
    """Continuous-batching DBA server over odba_server_* (slots 1..32). The
    slots' streams overlap only with one hardware queue each: the process
    must start HIP with GPU_MAX_HW_QUEUES >= slots (HIP's default is 4)."""

    def __init__(self, device=0, hbm_budget=0, slots=4):
        self.L = N.load()
        self.h = C.c_void_p()
        rc = self.L.odba_server_create(device, hbm_budget, slots, C.byref(self.h))
        if rc:
            raise N.OdbaError(rc, "odba_server_create failed")

    def submit(self, seqs, open_start=0, open_end=0, norm=1, max_rounds=250):
        seqs = [np.ascontiguousarray(s) for s in seqs]
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        jid = C.c_uint64()
        rc = self.L.odba_server_submit(self.h, N.CODE[seqs[0].dtype], N._ptrs(seqs), N._p(lens), len(seqs),
                                       open_start, open_end, norm, max_rounds, C.byref(jid))
        if rc:
            raise N.OdbaError(rc, "odba_server_submit failed")
        return jid.value, seqs[0].dtype, int(lens.max())

    def wait(self, jid, timeout=-1.0):
        st = C.c_int()
        rc = self.L.odba_server_wait(self.h, jid, timeout, C.byref(st))
        if rc:
            raise N.OdbaError(rc, "odba_server_wait failed")
        return st.value

    def result(self, jid, dtype, maxlen):
        cen = np.zeros(maxlen, dtype=dtype)
        clen = C.c_uint64()
        med, rounds, jrc = C.c_int(), C.c_int(), C.c_int()
        times = (C.c_double * 3)()
        rc = self.L.odba_server_result(self.h, jid, N._p(cen), maxlen, C.byref(clen), C.byref(med), C.byref(rounds),
                                       times, C.byref(jrc))
        if rc:
            raise N.OdbaError(rc, "odba_server_result failed")
        return dict(centroid=cen[: clen.value], medoid=med.value, rounds=rounds.value, rc=jrc.value,
                    t_submit=times[0], t_admit=times[1], t_done=times[2])

    def release(self, jid):
        self.L.odba_server_release(self.h, jid)

    def stats(self):
        use, bud = C.c_uint64(), C.c_uint64()
        run, q = C.c_int(), C.c_int()
        self.L.odba_server_stats(self.h, C.byref(use), C.byref(bud), C.byref(run), C.byref(q))
        return dict(bytes_in_use=use.value, budget=bud.value, running=run.value, queued=q.value)

    def close(self):
        if self.h:
            self.L.odba_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else float("nan")


def run_trace(jobs, mode="served", slots=4, hbm_budget=0, max_rounds=250, norm=1, progress=None):
    """Replay `jobs` (dicts with arrival time in seconds, seqs, open_end) in
    real time. mode "served": submit each job to a Server at its arrival
    time; "plain": run odba_perform_dba for one job after another. Returns
    per-job results and latencies (completion - arrival)."""
    out = [None] * len(jobs)
    t_start = time.perf_counter()
    if mode == "plain":
        with N.Context(0) as ctx:
            for k, jb in enumerate(jobs):
                delay = jb["arrival"] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                r = ctx.perform_dba(jb["seqs"], 0, jb["open_end"], norm=norm, max_rounds=max_rounds)
                done = time.perf_counter() - t_start
                out[k] = dict(centroid=r["centroid"], medoid=r["medoid"], rounds=r["rounds"],
                              arrival=jb["arrival"], done=done, latency=done - jb["arrival"])
                if progress:
                    progress(k, out[k])
    else:
        with Server(0, hbm_budget, slots) as srv:
            handles = [None] * len(jobs)

            def submitter():
                for k, jb in enumerate(jobs):
                    delay = jb["arrival"] - (time.perf_counter() - t_start)
                    if delay > 0:
                        time.sleep(delay)
                    handles[k] = srv.submit(jb["seqs"], 0, jb["open_end"], norm, max_rounds)

            th = threading.Thread(target=submitter)
            th.start()
            th.join()
            for k, jb in enumerate(jobs):
                jid, dt, ml = handles[k]
                srv.wait(jid)
                r = srv.result(jid, dt, ml)
                if r["rc"]:
                    raise N.OdbaError(r["rc"], f"job {k} failed")
                # server times are seconds since the server was created, which
                # precedes t_start by the server set-up; arrival is the submit
                # time in the same clock
                done = r["t_done"] - r["t_submit"] + jb["arrival"]
                out[k] = dict(centroid=r["centroid"], medoid=r["medoid"], rounds=r["rounds"], arrival=jb["arrival"],
                              done=done, latency=r["t_done"] - r["t_submit"], queued=r["t_admit"] - r["t_submit"])
                srv.release(jid)
                if progress:
                    progress(k, out[k])
    makespan = max(o["done"] for o in out) - min(o["arrival"] for o in out)
    lat = [o["latency"] for o in out]
    return dict(results=out, jobs=len(jobs), makespan_s=makespan, jobs_per_s=len(jobs) / makespan,
                p50_s=_pct(lat, 50), p99_s=_pct(lat, 99), mean_s=float(np.mean(lat)))
