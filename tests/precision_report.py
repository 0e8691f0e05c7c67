"""Precision report: shot records and gradients of the GPU path, with FP32 and
with FP16 checkpoint storage, against the FP64 oracle run of the same
propagator on the same inputs (observed data = the FP64 oracle's modeling of
the true model).  One JSON line per storage.

    python tests/precision_report.py --workload c2 --shots 1 [--nt N]
        [--freq 5 --filter 0]

Test infrastructure: it runs the CPU oracle (OpenMP over the host's cores) and
is meant for one-off statements at sizes the parity tests do not reach.  The
oracle's FFT filters (filter 0/1) are plain DFTs, O(n^2) per trace: keep
filtered reports to small grids.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [os.path.join(ROOT, "fwi-edu_amd"), os.path.join(ROOT, "oracle"), ROOT]

import numpy as np  # noqa: E402

from fwiedu import lib as flib  # noqa: E402
from fwiedu import synth  # noqa: E402
from fwiedu.records import Array2D  # noqa: E402
from oracle import Oracle  # noqa: E402


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--shots", type=int, default=1)
    ap.add_argument("--nt", type=int, default=0)
    ap.add_argument("--freq", type=float, default=0.0)
    ap.add_argument("--filter", type=int, default=2)
    args = ap.parse_args()
    over = {"nshot": args.shots, "freq": args.freq, "filter": args.filter}
    if args.nt:
        over["nt"] = args.nt
    w = synth.workload(args.workload, **over)
    L = flib.load()
    o64 = Oracle("f64")

    def log(msg):
        print(f"[{time.perf_counter() - t0:7.1f} s] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    log(f"{w.name}: {w.nz}x{w.nx} nt {w.nt}, {w.nshot} shot(s); FP64 oracle modeling")
    obs = w.shots()
    vpt, rht = Array2D(w.vp_true), Array2D(w.rho_true)
    o64.propagation(vpt.pp, rht.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                    w.freesurface, w.freq, float(w.filter), w.precondition, 0, obs.recs, None)
    log("FP64 oracle gradient")
    vps, rhs = Array2D(w.vp_start), Array2D(w.rho_start)
    pre64 = obs.header_copy()
    g64 = Array2D(np.zeros((w.nx, w.nz), np.float32))
    o64.propagation(vps.pp, rhs.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                    w.freesurface, w.freq, float(w.filter), w.precondition, 1, pre64.recs, g64.pp)
    t_cpu = time.perf_counter() - t0
    # GPU records of the true model
    rec_g = w.shots()
    flib.propagate(L, w, vpt, rht, rec_g, 0)
    rec_err = [rel(rec_g.trace(s), obs.trace(s)) for s in range(w.nshot)]
    for bits in (32, 16):
        log(f"GPU gradient, fp{bits} checkpoints")
        L.fwiedu_set_checkpoint_bits(bits)
        pre = obs.header_copy()
        g = Array2D(np.zeros((w.nx, w.nz), np.float32))
        t1 = time.perf_counter()
        flib.propagate(L, w, vps, rhs, obs, 1, pre, g)
        t_gpu = time.perf_counter() - t1
        L.fwiedu_set_checkpoint_bits(32)
        print(json.dumps({
            "workload": f"{w.name}: {w.nz}x{w.nx}, nt {w.nt}, {w.nshot} shot(s), filter {w.filter}, "
                        f"freq {w.freq}",
            "checkpoint_storage": f"fp{bits}",
            "records_rel_l2_vs_fp64_per_shot": rec_err,
            "adjoint_source_rel_l2_vs_fp64_per_shot":
                [rel(pre.trace(s), pre64.trace(s)) for s in range(w.nshot)],
            "gradient_rel_l2_vs_fp64": rel(g.a, g64.a),
            "gpu_gradient_s": t_gpu, "fp64_oracle_s": t_cpu}), flush=True)


if __name__ == "__main__":
    main()
