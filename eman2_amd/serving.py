"""Python side of the fourier_hip serving loop (fh_server_* in the C ABI).

``Server`` wraps the native scheduler (FIFO admission by HBM footprint, one
worker per active job, each job's volume resident on the device on its own
stream).  ``poisson_stream`` drives it with a seeded Poisson arrival process
and reports per-job latency (submit -> finished map on the host).
"""
import ctypes
import math
import time

import numpy as np

from . import _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a.ctypes.data_as(ctypes.c_void_p)


class Server:
    def __init__(self, device=0, hbm_budget_bytes=64 << 30, max_active=4, chunk=2048):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.fh_server_create(int(device), int(hbm_budget_bytes), int(max_active), int(chunk),
                                        ctypes.byref(h))
        if rc:
            raise _lib.FourierHipError(rc, "fh_server_create")
        self._h = h
        self._keep = {}

    def submit(self, size, imgs, xforms, weights=None, defocus=None, ctf=None, sym="c1", mode="gauss_2",
               box=None, out_map=True):
        """imgs: numpy (n,box,box) on the host, or a device pointer (int)
        with ``box``; xforms (n,3,4).  Returns the job id."""
        n = len(xforms)
        m = np.ascontiguousarray(np.asarray(xforms, np.float32).reshape(n, 12))
        w = np.ascontiguousarray(np.ones(n, np.float32) if weights is None else weights, np.float32)
        d = None if defocus is None else np.ascontiguousarray(defocus, np.float32)
        if isinstance(imgs, int):
            where, B, keep_img = _lib.FH_DEVICE, int(box), None
        else:
            keep_img = np.ascontiguousarray(imgs, np.float32)
            where, B = _lib.FH_HOST, keep_img.shape[-1]
        P = int(size)
        out = np.zeros((P, P, P), np.float32) if out_map else None
        prm = _lib.Params(P, P, P, sym.encode(), mode.encode(), 0, 0, 0, 0, 0)
        c = _lib.Ctf(*(float(ctf.get(k, dflt)) for k, dflt in (("voltage", 300.0), ("cs", 2.7), ("ampcont", 10.0),
                                                                ("apix", 1.0)))) if ctf else _lib.Ctf()
        job = _lib.Job(prm, _ptr(keep_img) if keep_img is not None else _ptr(imgs), n, B, where, _ptr(m), _ptr(w),
                       _ptr(d), c, int(ctf is not None and d is not None), _ptr(out))
        jid = ctypes.c_int()
        rc = self._lib.fh_server_submit(self._h, ctypes.byref(job), ctypes.byref(jid))
        if rc:
            raise _lib.FourierHipError(rc, self._lib.fh_server_last_error(self._h).decode())
        self._keep[jid.value] = (keep_img, m, w, d, out, prm)
        return jid.value

    def poll(self, jid):
        st = _lib.JobStatus()
        _lib.check(self._lib.fh_server_poll(self._h, int(jid), ctypes.byref(st)))
        return {"state": _lib.JOB_STATES[st.state], "error": st.error, "t_submit": st.t_submit,
                "t_admit": st.t_admit, "t_done": st.t_done, "particles_done": st.particles_done,
                "footprint_bytes": st.footprint_bytes}

    def result(self, jid):
        return self._keep[jid][4]

    def cancel(self, jid):
        self._lib.fh_server_cancel(self._h, int(jid))

    def drain(self):
        self._lib.fh_server_drain(self._h)

    def clock(self):
        return self._lib.fh_server_clock(self._h)

    def close(self):
        if self._h is not None:
            self._lib.fh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_stream(server, jobs, rate, seed):
    """Submit ``jobs`` (list of dicts of Server.submit kwargs) at Poisson
    arrival times (rate jobs/s), wait for all, return per-job records."""
    rng = np.random.default_rng([seed, 0x5E])
    gaps = rng.exponential(1.0 / rate, len(jobs))
    t0 = time.perf_counter()
    ids = []
    t_next = 0.0
    for j, g in zip(jobs, gaps):
        t_next += g
        dt = t_next - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(**j))
    server.drain()
    out = []
    for jid, j in zip(ids, jobs):
        st = server.poll(jid)
        st["latency_s"] = st["t_done"] - st["t_submit"]
        st["particles"] = len(j["xforms"])
        st["size"] = j["size"]
        out.append(st)
    return out


def latency_summary(records):
    lat = np.array([r["latency_s"] for r in records if r["state"] == "done"])
    if len(lat) == 0:
        return {}
    t_first = min(r["t_submit"] for r in records)
    t_last = max(r["t_done"] for r in records)
    parts = sum(r["particles"] for r in records if r["state"] == "done")
    return {"jobs": len(records), "done": int(len(lat)), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "mean_ms": round(float(lat.mean()) * 1e3, 1),
            "particles_per_s": round(parts / max(t_last - t_first, 1e-9), 1),
            "wall_s": round(t_last - t_first, 2)}


def config4_jobs(n_jobs, seed, pools, rng=None):
    """Serving workload: box in {64,128,192,256,320} -> pad 1.5x good size,
    particle counts log-uniform 500-20,000, sym in {c1,c4,d2}, CTF, SNR 0.1.
    ``pools`` maps box -> (device pointer or array, xforms, defocus, count)
    of pre-generated particles; each job takes a random contiguous run."""
    from .synthetic import good_size
    rng = rng or np.random.default_rng([seed, 4])
    jobs = []
    boxes = sorted(pools)
    for _ in range(n_jobs):
        B = int(rng.choice(boxes))
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(20000)))))
        src, mats, dfs, cap, apix, elem_ptr = pools[B]
        n = min(n, cap)
        o = int(rng.integers(0, cap - n + 1))
        imgs = elem_ptr(o) if callable(elem_ptr) else src[o:o + n]
        jobs.append(dict(size=good_size(int(round(1.5 * B))), imgs=imgs, xforms=mats[o:o + n],
                         defocus=dfs[o:o + n], ctf={"voltage": 300.0, "cs": 2.7, "ampcont": 10.0, "apix": apix},
                         sym=str(rng.choice(["c1", "c4", "d2"])), box=B))
    return jobs
