"""Serving loop on the GPU: results are the same alone and in the mix,
admission by HBM footprint, failure isolation, cancellation.
Bar: max|a-b|/max|b| <= 1e-4 and FSC >= 0.9999 between a job's map alone
and in the mix (FP32 atomics make summation order differ)."""
import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.metrics import fsc, max_rel_err

pytestmark = pytest.mark.gpu
CTF = {"voltage": 300.0, "cs": 2.7, "ampcont": 10.0, "apix": 1.0}


def make_jobs():
    specs = [(48, 32, 300, "c1"), (64, 40, 200, "c4"), (96, 64, 400, "d2"), (48, 32, 150, "c1")]
    jobs = []
    for i, (P, B, n, sym) in enumerate(specs):
        ph = S.phantom(12, B * 0.3, seed=100 + i)
        imgs, mats, dfs = S.particles(ph, n, B, seed=200 + i, snr=0.1, ctf={})
        jobs.append(dict(size=P, imgs=imgs, xforms=mats, defocus=dfs, ctf=CTF, sym=sym))
    return jobs


def test_alone_equals_mix():
    from eman2_amd.serving import Server
    jobs = make_jobs()
    alone = []
    s1 = Server(max_active=1, chunk=128)
    for j in jobs:
        jid = s1.submit(**j)
        s1.drain()
        assert s1.poll(jid)["state"] == "done"
        alone.append(s1.result(jid).copy())
    s1.close()
    s4 = Server(max_active=4, chunk=64)
    ids = [s4.submit(**j) for j in jobs]
    s4.drain()
    for jid, a in zip(ids, alone):
        st = s4.poll(jid)
        assert st["state"] == "done", st
        m = s4.result(jid)
        print("job", jid, max_rel_err(m, a), fsc(m, a).min(), flush=True)
        assert max_rel_err(m, a) <= 1e-4
        assert fsc(m, a).min() >= 0.9999
    s4.close()


def test_matches_direct_api():
    from eman2_amd import Reconstructors
    from eman2_amd.serving import Server
    j = make_jobs()[2]
    r = Reconstructors.get("fourier_hip", {"size": (j["size"],) * 3, "sym": j["sym"]})
    r.setup()
    r.insert_batch(j["imgs"], j["xforms"], None, defocus=j["defocus"], ctf=CTF)
    direct = r.finish(True, return_numpy=True)
    s = Server(max_active=2)
    jid = s.submit(**j)
    s.drain()
    assert max_rel_err(s.result(jid), direct) <= 1e-4
    s.close()


def test_admission_by_footprint_and_rejection():
    from eman2_amd.serving import Server
    jobs = make_jobs()
    probe = Server(max_active=1, chunk=64)  # same chunk as below: same footprint
    fp = [probe.poll(probe.submit(**j))["footprint_bytes"] for j in jobs[:1]]
    probe.drain()
    probe.close()
    budget = int(fp[0] * 1.5)  # fits one 48^3 job at a time, not two
    s = Server(hbm_budget_bytes=budget, max_active=4, chunk=64)
    a = s.submit(**jobs[0])
    b = s.submit(**jobs[3])
    big = s.submit(**jobs[2])  # a 96^3 job does not fit the budget at all
    s.drain()
    sa, sb, sbig = s.poll(a), s.poll(b), s.poll(big)
    assert sa["state"] == "done" and sb["state"] == "done"
    assert sb["t_admit"] >= sa["t_done"] - 1e-6  # waited for memory
    assert sbig["state"] == "rejected"
    s.close()


def test_bad_job_fails_alone_and_cancel():
    from eman2_amd.serving import Server
    jobs = make_jobs()
    s = Server(max_active=1, chunk=64)
    good = s.submit(**jobs[0])
    bad = s.submit(**dict(jobs[3], sym="q9"))
    queued = s.submit(**jobs[1])
    s.cancel(queued)
    s.drain()
    assert s.poll(good)["state"] == "done"
    st = s.poll(bad)
    assert st["state"] == "failed" and st["error"] == 2
    assert s.poll(queued)["state"] in ("cancelled", "done")
    s.close()


def test_served_job_other_symmetry_and_mode():
    """a served icos / gauss_var job is the direct API's result"""
    from eman2_amd import Reconstructors
    from eman2_amd.serving import Server
    j = dict(make_jobs()[0], sym="icos", mode="gauss_var")
    r = Reconstructors.get("fourier_hip", {"size": (j["size"],) * 3, "sym": "icos", "mode": "gauss_var"})
    r.setup()
    r.insert_batch(j["imgs"], j["xforms"], None, defocus=j["defocus"], ctf=CTF)
    direct = r.finish(True, return_numpy=True)
    s = Server(max_active=2)
    jid = s.submit(**j)
    s.drain()
    assert s.poll(jid)["state"] == "done"
    m = s.result(jid)
    assert max_rel_err(m, direct) <= 1e-4 and fsc(m, direct).min() >= 0.9999
    s.close()


@pytest.mark.parametrize("sym,nsym", [("icos", 60), ("oct", 24), ("tet", 12), ("d7", 14), ("c4", 4)])
def test_footprint_counts_every_symmetry_operator(sym, nsym):
    """a job's admission footprint is the library's footprint for the number
    of operators its symmetry really has, plus host staging and finish()"""
    from eman2_amd import _lib
    from eman2_amd.serving import Server
    j = dict(make_jobs()[0], sym=sym)
    chunk = 64
    s = Server(max_active=1, chunk=chunk)
    jid = s.submit(**j)
    fp = s.poll(jid)["footprint_bytes"]
    s.cancel(jid)
    s.drain()
    s.close()
    P, B, m = j["size"], j["imgs"].shape[-1], min(chunk, len(j["xforms"]))
    L = _lib.load()
    assert L.fh_get_symmetries(sym.encode(), None, 0) == nsym
    want = L.fh_footprint_bytes(P, nsym, m) + 4 * m * B * B + 4 * P ** 3 + 8 * P * P
    assert fp == want, (fp, want)


def test_alone_equals_mix_with_pad480_job():
    """a box-320 / pad-480 job of the serving mix (compile-time FFT plans of
    pads 192 and 480) gives the same map alone and beside smaller jobs"""
    from eman2_amd.serving import Server
    specs = [(480, 320, 120, "d2"), (192, 128, 300, "c4"), (96, 64, 400, "c1")]
    jobs = []
    for i, (P, B, n, sym) in enumerate(specs):
        ph = S.phantom(30, B * 0.3, seed=300 + i)
        imgs, mats, dfs = S.particles(ph, n, B, seed=400 + i, snr=0.1, ctf={})
        jobs.append(dict(size=P, imgs=imgs, xforms=mats, defocus=dfs, ctf=CTF, sym=sym))
    alone = []
    s1 = Server(max_active=1, chunk=128)
    for j in jobs:
        jid = s1.submit(**j)
        s1.drain()
        assert s1.poll(jid)["state"] == "done"
        alone.append(s1.result(jid).copy())
    s1.close()
    s3 = Server(max_active=3, chunk=64)
    ids = [s3.submit(**j) for j in jobs]
    s3.drain()
    for jid, a in zip(ids, alone):
        assert s3.poll(jid)["state"] == "done"
        m = s3.result(jid)
        print("job", jid, max_rel_err(m, a), fsc(m, a).min(), flush=True)
        assert max_rel_err(m, a) <= 1e-4 and fsc(m, a).min() >= 0.9999
    s3.close()
