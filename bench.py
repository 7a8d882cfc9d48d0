#!/usr/bin/env python3
"""Benchmark: particles inserted/s at 256^3 box (pad 384), c1, gauss_2.

Workload (one step = one reconstruction pass over the whole batch):
  50,000 float32 projections, 256 px box, 1.0 A/px, of a seeded 200-Gaussian
  phantom at uniform SO(3) orientations, CTF (300 kV, Cs 2.7 mm, 10% amplitude
  contrast, defocus U(0.8,2.5) um, B 80 A^2), SNR 0.05 noise; CTF phase flip,
  edge-mean subtraction, zero-pad to 384, r2c FFT and gauss_2 gridding into a
  384^3 (complex, weight) volume -- fh_insert_batch with the particles already
  resident in HBM.  Each step re-runs setup() and inserts all particles.

Prints one JSON line with the throughput, the brick scatter insertion
kernel's roofline fraction (HBM-bound scatter model), the plain one-particle-per-launch
GPU form, and the CPU oracle timed on this host's cores on a bounded sample.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
ENGINE_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
LDS_RANDOM_RMW_PER_CLK = 4.2  # plain LDS read-modify-writes per clock per CU at random addresses (microbenchmark)
CTF = dict(voltage=300.0, cs=2.7, ampcont=10.0, bfactor=80.0)


def npix(P):
    h = P // 2
    y = np.arange(-h, h)[:, None]
    x = np.arange(0, h + 1)[None, :]
    return int(((x * x + y * y) <= h * h + h).sum())


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=50000)
    ap.add_argument("--box", type=int, default=256)
    ap.add_argument("--pad", type=int, default=384)
    ap.add_argument("--sym", default="c1")
    ap.add_argument("--mode", default="gauss_2")
    ap.add_argument("--apix", type=float, default=1.0)
    ap.add_argument("--snr", type=float, default=0.05)
    ap.add_argument("--wpb", type=int, default=None, help="waves (item workers) per workgroup")
    ap.add_argument("--spi", type=int, default=None, help="slices per work item")
    ap.add_argument("--chunk", type=int, default=None)
    ap.add_argument("--bpc", type=int, default=None, help="persistent workgroups per CU")
    ap.add_argument("--graph", type=int, default=None, help="1: each chunk as one captured hipGraph")
    ap.add_argument("--plain-sample", type=int, default=200)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-sample", type=int, default=4096, help="particles copied to the host for the CPU leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB elementwise copy (PMC byte-count calibration)")
    ap.add_argument("--serve-jobs", type=int, default=40,
                    help="jobs in the serving measurement (0 = skip); Poisson arrivals at --serve-rate")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other measured shapes (100,000 x 192 px d7 / pad 288; the refine flow at pad 600)")
    ap.add_argument("--refine-particles", type=int, default=50000,
                    help="particles of the refine-flow line measured after the main one")
    ap.add_argument("--workload", default="insert", choices=["insert", "refine"],
                    help="insert: batched insertion (the metric); refine: insert, slice agreement of every "
                         "particle, re-insertion of img x reconstruct_norm, finish() (400 px -> 600 by default)")
    return ap.parse_args()


def serving(a, torch, dev):
    """Mixed-size job stream through the native serving loop: box in
    {64,128,192,256,320} (pad 1.5x to a good size), 500-20,000 particles
    (log-uniform), sym in {c1,c4,d2}, CTF, SNR 0.1, Poisson arrivals.
    Particles come from per-box pools generated in HBM beforehand; each job
    takes a random contiguous run of its box's pool."""
    from eman2_amd import serving as SV
    from eman2_amd import synthetic as S

    pools = {}
    for i, B in enumerate((64, 128, 192, 256, 320)):
        cap = 20000
        ph = S.phantom(60, 0.3 * B, seed=40 + i)
        imgs, mats, dfs = S.particles_device(ph, cap, B, seed=50 + i, apix=1.0, snr=0.1, ctf=CTF, device=dev)
        base = int(imgs.data_ptr())
        pools[B] = (imgs, mats, dfs, cap, 1.0, (lambda o, base=base, B=B: base + 4 * B * B * o))
    torch.cuda.synchronize()
    jobs = SV.config4_jobs(a.serve_jobs, seed=9, pools=pools)
    srv = SV.Server(device=0, hbm_budget_bytes=160 << 30, max_active=4, chunk=1024)
    # warm the code paths once per size
    for B in sorted(pools):
        j = dict(SV.config4_jobs(1, seed=B, pools={B: pools[B]})[0])
        j["xforms"], j["defocus"] = j["xforms"][:64], j["defocus"][:64]
        srv.submit(**j)
    srv.drain()
    recs = SV.poisson_stream(srv, jobs, a.serve_rate, seed=11)
    srv.close()
    out = SV.latency_summary(recs)
    out.update({"rate_jobs_per_s": a.serve_rate, "max_active": 4,
                "what": "latency = submit -> finished map on the host (setup, insertion, finish, D2H)"})
    del pools
    torch.cuda.empty_cache()
    return out


def cpu_baseline(imgs_host, mats, dfs, P, sym, mode, apix, budget_s):
    """Oracle (-O3 -ffast-math build) on this host's cores: prep (edge mean,
    pad, CTF flip, r2c) in a thread pool + multi-threaded insertion with
    private volumes.  Sample sized for ~budget_s of CPU work."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    L = O.lib(fast=True)
    # every core this process may run on, within the host's CPU share per GPU
    # (OMP_NUM_THREADS, which the GPU host sets to its share)
    aff = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    T = max(1, min(aff, share) if share > 0 else aff)
    B = imgs_host.shape[-1]

    def prep(i):
        out = np.zeros((P, P + 2), np.float32)
        c = O.ctf_struct(dict(CTF, apix=apix, defocus=float(dfs[i]))) if dfs is not None else None
        raw = np.ascontiguousarray(imgs_host[i])
        L.or_prep_particle(O._f(raw), B, P, ctypes.byref(c) if c is not None else None, O._f(out))
        return out

    def run(k):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(T) as ex:
            sl = np.stack(list(ex.map(prep, range(k))))
        tp = time.perf_counter() - t0
        _, _, ti = O.insert_batch_mt(P, sl, mats[:k], np.ones(k, np.float32), sym, mode, nthreads=T)
        return tp, ti

    # size the sample: two probes separate the per-particle cost from the
    # fixed cost of the private volumes (allocation and the final sum)
    k1 = min(len(imgs_host), 2 * T)
    k2 = min(len(imgs_host), 4 * T)
    p1, i1 = run(k1)
    p2, i2 = run(k2)
    per = max((p2 + i2 - p1 - i1) / max(1, k2 - k1), (p2 + i2) / k2 * 0.05)
    n = int(min(len(imgs_host), max(k2, budget_s / per)))
    n = max(T, (n // T) * T)
    t_prep, t_ins = run(n)
    # one thread: prep and insertion of a small sample in this thread
    k1 = min(len(imgs_host), 16)
    t0 = time.perf_counter()
    sl = np.stack([prep(i) for i in range(k1)])
    tp1 = time.perf_counter() - t0
    _, _, ti1 = O.insert_batch_mt(P, sl, mats[:k1], np.ones(k1, np.float32), sym, mode, nthreads=1)
    return {
        "value": n / (t_prep + t_ins), "unit": "particles/s", "cores": T, "kind": "port",
        "affinity_cores": aff, "cpu_share": share or None,
        "sample": "%d particles of the same workload (prep %.2fs + insertion %.2fs, %d threads, "
                  "oracle -O3 -ffast-math, private volumes summed)" % (n, t_prep, t_ins, T),
        "prep_s": t_prep, "insert_s": t_ins, "insert_only_particles_per_s": n / t_ins,
        "single_thread": {"value": k1 / (tp1 + ti1), "unit": "particles/s", "sample": k1,
                          "prep_s": tp1, "insert_s": ti1},
    }


def refine(a, quiet=False):
    """Two-pass refinement flow (e2make3d.py:391-539): insert all particles,
    determine_slice_agreement(sub=True) for every particle, clear, re-insert
    each image multiplied by its reconstruct_norm, finish() with the 3-D
    inverse FFT.  Particles HBM-resident; the finished map is copied to the
    host as finish() returns it."""
    import torch

    from eman2_amd import Reconstructors
    from eman2_amd import synthetic as S
    from eman2_amd.autotune import nsym_of

    dev = "cuda:0"
    P, B, n = a.pad, a.box, a.particles
    print("[bench] refine flow: generating %d particles %dpx on %s" % (n, B, dev), file=sys.stderr, flush=True)
    ph = S.phantom(400, 96.0 * B / 256.0, seed=2)
    imgs, mats, dfs = S.particles_device(ph, n, B, seed=3, apix=a.apix, snr=a.snr, ctf=CTF, device=dev)
    torch.cuda.synchronize()
    w = np.ones(n, np.float32)
    ctf = dict(CTF, apix=a.apix)
    r = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": a.sym, "mode": a.mode})
    ptr = int(imgs.data_ptr())
    stages = {"insert1_ms": [], "agree_ms": [], "insert2_ms": [], "finish_ms": []}

    def step(record):
        t = [time.perf_counter()]
        r.setup()
        r.insert_batch(ptr, mats, w, defocus=dfs, ctf=ctf, box=B)
        r.sync()
        t.append(time.perf_counter())
        q = torch.empty((n, 5), dtype=torch.float32, device=dev)
        r.agreement_batch(ptr, mats, w, defocus=dfs, ctf=ctf, sub=True, apix=a.apix, box=B, out=int(q.data_ptr()))
        t.append(time.perf_counter())
        scaled = imgs * q[:, 0].view(-1, 1, 1)
        r.setup()
        r.insert_batch(int(scaled.data_ptr()), mats, w, defocus=dfs, ctf=ctf, box=B)
        r.sync()
        t.append(time.perf_counter())
        r.finish(True, return_numpy=True)
        t.append(time.perf_counter())
        del scaled
        if record:
            for k, i in zip(stages, range(4)):
                stages[k].append((t[i + 1] - t[i]) * 1e3)
        return t[-1] - t[0]

    for i in range(a.warmup):
        print("[bench] refine warmup %d" % i, file=sys.stderr, flush=True)
        step(False)
    total = 0.0
    for i in range(a.steps):
        total += step(True)
        print("[bench] refine step %d %s" % (i, {k: round(v[-1], 1) for k, v in stages.items()}), file=sys.stderr,
              flush=True)
    value = n * a.steps / total
    nsym = nsym_of(a.sym)
    H = P // 2
    # algorithmic bytes: two insertions (raw read + 12 B per gridding contribution),
    # the agreement pass (two scatters, the compare gather) and finish (normalise, 3-D FFT, map)
    b_ins = 4 * B * B + nsym * npix(P) * 96
    b_agree = 4 * B * B + 2 * nsym * npix(P) * 96 + 12 * (H + 2) * P * nsym
    b_fin = 20 * (H + 1) * P * P + 3 * 16 * (H + 1) * P * P + 4 * P ** 3
    job_bytes = n * (2 * b_ins + b_agree) + b_fin
    ag = sum(stages["agree_ms"]) / a.steps
    out = {
        "metric": "particles/s through the refine flow (insert, slice agreement, re-insert, finish)",
        "value": round(value, 1), "unit": "particles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(total / a.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d x %dpx particles, pad %d, %s, %s, CTF phase flip, SNR %g, HBM-resident" % (
            n, B, P, a.sym, a.mode, a.snr), "box": B, "pad": P, "particles": n, "sym": a.sym, "mode": a.mode},
        "stage_ms_per_step": {k: round(sum(v) / a.steps, 2) for k, v in stages.items()},
        "agreement_particles_per_s": round(n / (ag * 1e-3), 1),
        "pipeline_roofline": {"effective_GBs": round(job_bytes * a.steps / total / 1e9, 1),
                              "frac": round(job_bytes * a.steps / total / 1e9 / HBM_PEAK_GBS, 4),
                              "bytes_per_job": job_bytes},
    }
    del imgs
    torch.cuda.empty_cache()
    if not quiet:
        print(json.dumps(out), flush=True)
    return out


def measure_insert(P, B, n, sym, mode, steps, warmup, apix=1.0, snr=0.05, phantom=None):
    """batched insertion throughput of one shape (particles HBM-resident,
    CTF phase flip): setup() + insert_batch of all n particles per step"""
    import torch

    from eman2_amd import Reconstructors
    from eman2_amd import synthetic as S

    dev = "cuda:0"
    print("[bench] shape %d x %dpx pad %d %s %s" % (n, B, P, sym, mode), file=sys.stderr, flush=True)
    ph = phantom if phantom is not None else S.phantom(200, 96.0 * B / 256.0, seed=2)
    imgs, mats, dfs = S.particles_device(ph, n, B, seed=4, apix=apix, snr=snr, ctf=CTF, device=dev)
    torch.cuda.synchronize()
    w = np.ones(n, np.float32)
    ctf = dict(CTF, apix=apix)
    r = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": sym, "mode": mode})
    ptr = int(imgs.data_ptr())
    for _ in range(warmup):
        r.setup()
        r.insert_batch(ptr, mats, w, defocus=dfs, ctf=ctf, box=B)
    r.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.setup()
        r.insert_batch(ptr, mats, w, defocus=dfs, ctf=ctf, box=B)
    r.sync()
    el = time.perf_counter() - t0
    st = r.get_stats()
    del imgs, r
    torch.cuda.empty_cache()
    return {"value": round(steps * n / el, 1), "unit": "particles/s", "ms_per_step": round(el / steps * 1e3, 2),
            "steps": steps, "warmup": warmup, "stage_ms_last_step": {k: round(st[k], 2) for k in (
                "prep_ms", "bin_ms", "insert_kernel_ms", "merge_ms", "total_ms")},
            "config": "%d x %dpx particles, pad %d, %s, %s, CTF phase flip, HBM-resident" % (n, B, P, sym, mode)}


def main():
    a = parse()
    if a.workload == "refine":
        if a.gpus != 1:
            print("bench.py: measures one GPU only; --gpus %d is not supported" % a.gpus, file=sys.stderr)
            sys.exit(2)
        return refine(a)
    if a.gpus != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py: measures one GPU only (the path shards by particle, one volume replica per GPU "
                  "summed by an all-reduce); "
                  "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch

    from eman2_amd import Reconstructors
    from eman2_amd import synthetic as S
    from eman2_amd.autotune import nsym_of

    dev = "cuda:0"
    P, B, n = a.pad, a.box, a.particles
    nsym = nsym_of(a.sym)
    print("[bench] generating %d particles %dpx on %s" % (n, B, dev), file=sys.stderr, flush=True)
    ph = S.phantom(200, 96.0 * B / 256.0, seed=2)
    imgs, mats, dfs = S.particles_device(ph, n, B, seed=3, apix=a.apix, snr=a.snr, ctf=CTF, device=dev)
    torch.cuda.synchronize()
    weights = np.ones(n, np.float32)
    ctf = dict(CTF, apix=a.apix)

    r = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": a.sym, "mode": a.mode})
    tune = r.get_tuning()
    for k, v in (("waves_per_block", a.wpb), ("slices_per_item", a.spi), ("chunk", a.chunk), ("blocks_per_cu", a.bpc),
                 ("graph", a.graph)):
        if v is not None:
            tune[k] = v
    r.set_tuning(**tune)
    ptr = int(imgs.data_ptr())

    def step():
        r.setup()
        r.insert_batch(ptr, mats, weights, defocus=dfs, ctf=ctf, box=B)

    for i in range(a.warmup):
        print("[bench] warmup %d" % i, file=sys.stderr, flush=True)
        step()
    torch.cuda.synchronize()
    r.sync()
    lat, kern_ms, stats = [], 0.0, []
    t0 = time.perf_counter()
    for i in range(a.steps):
        ts = time.perf_counter()
        step()
        lat.append((time.perf_counter() - ts) * 1e3)
        st = r.get_stats()
        stats.append(st)
        kern_ms += st["insert_kernel_ms"]
        print("[bench] step %d %.1f ms (prep %.1f bin %.1f insert %.1f merge %.1f)" % (
            i, lat[-1], st["prep_ms"], st["bin_ms"], st["insert_kernel_ms"], st["merge_ms"]), file=sys.stderr, flush=True)
    r.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = a.steps * n / elapsed

    # roofline of the dominant kernel (k_scatter_insert; k_scatter_wide for
    # the wide stencils): scatter model, one 12-byte (complex + weight) update
    # per gridding contribution, 8 per pixel per symmetry copy (S^3 for a
    # stencil of S cells per axis)
    wide = {"gauss_var": 6, "gauss_5": 6, "gauss_3": 3, "kaiser_bessel": 5, "gridding_5": 5, "gridding_7": 7,
            "kaiser_bessel_derived": 9}.get(a.mode)
    kname = "k_scatter_wide" if wide else "k_scatter_insert"
    bytes_per_particle = nsym * npix(P) * (wide ** 3 if wide else 8) * 12
    launches = sum(s["chunks"] for s in stats)
    avg_ms = kern_ms / max(1, launches)
    per_launch_bytes = bytes_per_particle * (a.steps * n) / max(1, launches)
    # the graph form does not split per-kernel times: no kernel figure then
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    traffic = None
    # HBM bytes per particle of the same kernel from the FETCH_SIZE /
    # WRITE_SIZE runs (scaled by a 1 GiB copy), when measured on this workload
    tf = os.path.join(ROOT, "profiles", "scatter_insert_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if t.get("P") == P and t.get("sym") == a.sym and t.get("box") == B and t.get("mode", "gauss_2") == a.mode:
            traffic = round(t["hbm_bytes_per_particle"] * (a.steps * n) / max(1, launches))
    roof = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic,
            "kernel": kname, "avg_launch_ms": round(avg_ms, 3),
            "algorithmic_bytes_per_launch": int(per_launch_bytes)}
    if wide and achieved:
        # the wide-stencil scatter accumulates in LDS and is VALU-issue bound: its roofline is the
        # CUs' VALU issue rate (one wave64 instruction per CU per clock) over the VALU wave
        # instructions per contribution that the PMC run of this shape counted
        contrib = nsym * npix(P) * wide ** 3 * (a.steps * n) / max(1, launches)
        gcs = contrib / (avg_ms * 1e-3) / 1e9
        peak = None
        pf = os.path.join(ROOT, "profiles", "scatter_wide_pmc.json")
        if os.path.exists(pf):
            with open(pf) as f:
                t = json.load(f)
            if t.get("P") == P and t.get("sym") == a.sym and t.get("box") == B and t.get("mode") == a.mode:
                cus = torch.cuda.get_device_properties(0).multi_processor_count
                peak = cus * ENGINE_CLOCK_HZ / t["valu_insts_per_contribution"] / 1e9
        roof = {"bound": "valu", "achieved": round(gcs, 1), "peak": round(peak, 1) if peak else None,
                "unit": "G contributions/s", "frac": round(gcs / peak, 4) if peak else None, "traffic": None,
                "kernel": kname, "avg_launch_ms": round(avg_ms, 3), "contributions_per_launch": int(contrib),
                "scatter_model_GBs": round(achieved, 1)}
    # the gauss_2 / trilinear scatter's own limiter is the LDS: three plain
    # read-modify-writes (re, im, weight) per contribution against the
    # random-address read-modify-write rate of one CU's LDS that a
    # microbenchmark on this GPU measured (4.2 per clock)
    lds = None
    if not wide and avg_ms > 0:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        rmw = 3 * nsym * npix(P) * 8 * (a.steps * n) / max(1, launches)
        rate = rmw / (avg_ms * 1e-3) / (cus * ENGINE_CLOCK_HZ)
        lds = {"rmw_per_clk_per_cu": round(rate, 3), "random_rmw_ceiling": LDS_RANDOM_RMW_PER_CLK,
               "frac": round(rate / LDS_RANDOM_RMW_PER_CLK, 4), "rmw_per_launch": int(rmw)}
    # whole-pipeline effective bandwidth: the scatter-model bytes plus the 4*B^2 raw read
    pipe_bytes = (4 * B * B + bytes_per_particle) * a.steps * n
    pipeline = {"effective_GBs": round(pipe_bytes / elapsed / 1e9, 1),
                "frac": round(pipe_bytes / elapsed / 1e9 / HBM_PEAK_GBS, 4)}
    # strict compulsory traffic: the raw particle read once per particle, and
    # finish() (normalise, 3-pass 3D FFT, real output) once per job
    H = P // 2
    finish_bytes = 20 * (H + 1) * P * P + 3 * 16 * (H + 1) * P * P + 4 * P ** 3
    bound = HBM_PEAK_GBS * 1e9 / (4 * B * B)
    compulsory = {"bytes_per_particle": 4 * B * B, "finish_bytes_per_job": finish_bytes,
                  "particles_per_s_at_peak": round(bound, 1), "frac": round(value / bound, 4)}

    out = {
        "metric": "particles inserted/s @256³ c1 gauss_2 (batched vs plain); p50/p99 job latency",
        "value": round(value, 1), "unit": "particles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d x %dpx particles, pad %d, %s, %s, CTF phase flip, SNR %g, HBM-resident" % (
            n, B, P, a.sym, a.mode, a.snr), "box": B, "pad": P, "particles": n, "sym": a.sym, "mode": a.mode,
            "tuning": r.get_tuning()},
        "roofline": roof,
        "scatter_lds": lds,
        "pipeline_roofline": pipeline,
        "compulsory_bound": compulsory,
        "job_latency_ms": {"p50": round(statistics.median(lat), 2),
                           "p99": round(sorted(lat)[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)], 2),
                           "what": "setup + insert_batch of all particles, per step"},
        "stage_ms_per_step": {k: round(sum(s[k] for s in stats) / a.steps, 2)
                              for k in ("prep_ms", "bin_ms", "insert_kernel_ms", "merge_ms", "insert_ms", "total_ms")},
        "pairs_per_step": int(sum(s["pairs"] for s in stats) / a.steps),
        "items_per_step": int(sum(s["items"] for s in stats) / a.steps),
    }

    # plain one-particle-per-launch form on a host-side sample
    host = None
    if not a.no_plain or not a.no_cpu:
        m = min(n, max(a.plain_sample, 0 if a.no_cpu else a.cpu_max_sample))
        host = imgs[:m].cpu().numpy()
    if not a.no_plain:
        print("[bench] plain form on %d particles" % a.plain_sample, file=sys.stderr, flush=True)
        rp = Reconstructors.get("fourier_hip", {"size": (P, P, P), "sym": a.sym, "mode": a.mode})
        rp.setup()
        k = a.plain_sample
        rp.insert_plain(host[:2], mats[:2], weights[:2], defocus=dfs[:2], ctf=ctf)  # warm
        rp.setup()
        t0 = time.perf_counter()
        rp.insert_plain(host[:k], mats[:k], weights[:k], defocus=dfs[:k], ctf=ctf)
        tp = time.perf_counter() - t0
        pv = k / tp
        out["plain"] = {"value": round(pv, 1), "unit": "particles/s", "sample": k,
                        "form": "per particle: host->device copy, prep launches, one insertion launch per "
                                "symmetry op with global atomics, synchronise after each"}
        out["batched_vs_plain"] = round(value / pv, 2)
        # the same per-particle launches with the particles already in HBM:
        # launch and synchronisation cost without the PCIe copies
        rp.setup()
        t0 = time.perf_counter()
        rp.insert_plain(ptr, mats[:k], weights[:k], defocus=dfs[:k], ctf=ctf, box=B)
        tpd = time.perf_counter() - t0
        out["plain_device"] = {"value": round(k / tpd, 1), "unit": "particles/s", "sample": k,
                               "form": "as plain, particles HBM-resident (no per-particle copy)"}
        out["batched_vs_plain_device"] = round(value / (k / tpd), 2)
        del rp
    del r
    if a.serve_jobs > 0:
        print("[bench] serving %d jobs" % a.serve_jobs, file=sys.stderr, flush=True)
        del imgs
        torch.cuda.empty_cache()
        out["serving"] = serving(a, torch, dev)
        out["job_latency_ms"] = {"p50": out["serving"].get("p50_ms"), "p99": out["serving"].get("p99_ms"),
                                 "what": "serving stream of mixed-size jobs (see serving)"}
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(host, mats, dfs, P, a.sym, a.mode, a.apix, a.cpu_seconds)
    if not a.no_configs:
        # the other shapes, measured in this run: the d7 shape (one subunit of
        # 30 Gaussians replicated by the 14 operators) and the two-pass refine flow
        from eman2_amd import _lib as LIB
        from eman2_amd import synthetic as S
        if "imgs" in locals():
            del imgs
        torch.cuda.empty_cache()
        ops = np.zeros((14, 9), np.float32)
        LIB.load().fh_get_symmetries(b"d7", ops.ctypes.data, 14)
        d7 = np.zeros((14, 3, 4), np.float32)
        d7[:, :, :3] = ops.reshape(14, 3, 3)
        ph3 = S.symmetric_phantom(S.phantom(30, 60.0, seed=5), d7)
        out["configs"] = {"d7_pad288": measure_insert(288, 192, 100000, "d7", "gauss_2", 2, 1, phantom=ph3)}
        # gauss_var, e2make3dpar's default inserter (6^3 stencils, the wide-stencil brick scatter)
        out["configs"]["gauss_var_pad384"] = measure_insert(384, 256, 50000, "c1", "gauss_var", 2, 1)
        ra = argparse.Namespace(**dict(vars(a), box=400, pad=600, apix=0.8, particles=a.refine_particles,
                                       steps=2, warmup=1, sym="c1", mode="gauss_2"))
        rf = refine(ra, quiet=True)
        out["configs"]["refine_pad600"] = {k: rf[k] for k in ("value", "unit", "ms_per_step", "config",
                                                               "stage_ms_per_step", "agreement_particles_per_s",
                                                               "pipeline_roofline")}
    if a.calibrate_copy:
        x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = {"read": x.numel() * 4, "write": y.numel() * 4}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
