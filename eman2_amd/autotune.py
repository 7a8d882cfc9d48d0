"""Autotuner for the batched insertion: sweeps waves per workgroup, slices
per work item, persistent workgroups per CU and particles per launch on a short
HBM-resident synthetic workload and records the fastest setting per
(pad, symmetry-op count) in tuning_table.json next to this file.

    python -m eman2_amd.autotune --box 256 --pad 384 --sym c1 --particles 50000

``FourierHipReconstructor`` applies the table entry whose key matches, or the
nearest pad with the same symmetry-op count, when it is created.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning_table.json")


def nsym_of(sym):
    """Number of symmetry operators of a symmetry name (Symmetry3D::get_nsym)."""
    from . import _lib
    n = _lib.load().fh_get_symmetries(str(sym).encode(), None, 0)
    if n < 0:
        raise ValueError("unknown symmetry %r" % (sym,))
    return n


def load_table():
    if not os.path.exists(TABLE):
        return {}
    with open(TABLE) as f:
        return json.load(f)


def lookup(P, nsym):
    """Best known tuning for pad P and nsym symmetry ops (or None)."""
    t = load_table()
    key = "%d,%d" % (P, nsym)
    if key in t:
        return t[key]["tuning"]
    cands = [(abs(int(k.split(",")[0]) - P), v["tuning"]) for k, v in t.items() if int(k.split(",")[1]) == nsym]
    if not cands:
        cands = [(abs(int(k.split(",")[0]) - P), v["tuning"]) for k, v in t.items()]
    return min(cands, key=lambda c: c[0])[1] if cands else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--box", type=int, default=256)
    ap.add_argument("--pad", type=int, default=384)
    ap.add_argument("--sym", default="c1")
    ap.add_argument("--particles", type=int, default=50000)
    ap.add_argument("--wpb", default="1,2,3,4")
    ap.add_argument("--spi", default="256,512,1024,2048")
    ap.add_argument("--bpc", default="8")
    ap.add_argument("--chunks", default="0,25000", help="0 = automatic (48 GiB of staging)")
    ap.add_argument("--graphs", default="0,1", help="1 = each chunk as one captured hipGraph")
    ap.add_argument("--write", action="store_true", help="update tuning_table.json")
    a = ap.parse_args()
    import torch

    from . import synthetic as S
    from .reconstructor import TUNING_KEYS, Reconstructors

    dev = "cuda:0"
    ph = S.phantom(200, 0.375 * a.box, seed=2)
    imgs, mats, dfs = S.particles_device(ph, a.particles, a.box, seed=3, snr=0.05,
                                         ctf={"voltage": 300.0, "cs": 2.7, "ampcont": 10.0}, device=dev)
    torch.cuda.synchronize()
    ctf = {"voltage": 300.0, "cs": 2.7, "ampcont": 10.0, "apix": 1.0}
    w = np.ones(a.particles, np.float32)
    r = Reconstructors.get("fourier_hip", {"size": (a.pad,) * 3, "sym": a.sym})
    ptr = int(imgs.data_ptr())
    results = []
    grid = itertools.product([int(x) for x in a.wpb.split(",")], [int(x) for x in a.spi.split(",")],
                             [int(x) for x in a.bpc.split(",")], [int(x) for x in a.chunks.split(",")],
                             [int(x) for x in a.graphs.split(",")])
    for wpb, spi, bpc, chunk, graph in grid:
        r.set_tuning(waves_per_block=wpb, slices_per_item=spi, chunk=chunk, blocks_per_cu=bpc, graph=graph)
        r.setup()
        r.insert_batch(ptr, mats, w, defocus=dfs, ctf=ctf, box=a.box)  # warm
        r.setup()
        t0 = time.perf_counter()
        r.insert_batch(ptr, mats, w, defocus=dfs, ctf=ctf, box=a.box)
        r.sync()
        dt = time.perf_counter() - t0
        st = r.get_stats()
        rec = {"waves_per_block": wpb, "slices_per_item": spi, "blocks_per_cu": bpc, "chunk": chunk, "graph": graph,
               "particles_per_s": a.particles / dt, "insert_kernel_ms_per_particle": st["insert_kernel_ms"] / a.particles}
        results.append(rec)
        print(json.dumps(rec), flush=True)
    best = max(results, key=lambda x: x["particles_per_s"])
    print("best", json.dumps(best), flush=True)
    if a.write:
        t = load_table()
        t["%d,%d" % (a.pad, nsym_of(a.sym))] = {
            "tuning": {k: best[k] for k in TUNING_KEYS},
            "particles_per_s": round(best["particles_per_s"], 1), "box": a.box, "particles": a.particles}
        with open(TABLE, "w") as f:
            json.dump(t, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    sys.exit(main())
