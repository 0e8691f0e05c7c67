#!/usr/bin/env python3
"""Benchmark of the FWI gradient hot path on one MI355X.

One step = one FWI gradient (iso_acoustic2d_propagation problem 1 semantics:
forward with checkpoints and pseudo-Hessian, adjoint source, reconstruction,
adjoint and imaging, per-shot and global preconditioning) over all shots of the
221 x 601 demo-shape workload (nt 4001, 120 shots x 601 receivers, pml 40, free
surface), with the model and observed records resident in HBM (a
fwiedu_session).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--plain-shots 4] [--cpu-shots 4]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "fwi-edu_amd"), os.path.join(ROOT, "oracle"), ROOT]

import numpy as np  # noqa: E402

from fwiedu import lib as flib  # noqa: E402
from fwiedu import synth  # noqa: E402
from fwiedu.records import Array2D  # noqa: E402

METRIC = "gradient shots/s at 221x601 (plain vs batched), % HBM roofline, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def interior_cells(w):
    Nz, Nx = w.nz + 2 * w.pml, w.nx + 2 * w.pml
    zs = w.pml if w.freesurface else 3
    return (Nx - 6) * (Nz - 3 - zs)


def observed(L, w):
    obs = w.shots()
    flib.propagate(L, w, Array2D(w.vp_true), Array2D(w.rho_true), obs, 0)
    return obs


def timed_gradients(L, w, obs, steps, warmup, batch, ktime_every=0):
    L.fwiedu_set_batch(batch)
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    sess = L.fwiedu_session_create(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt,
                                   w.nt, w.pml, w.freesurface, w.freq, float(w.filter),
                                   w.precondition)
    for i in range(warmup):
        L.fwiedu_session_gradient(sess, None)
        log(f"warmup {i + 1}/{warmup} done")
    L.fwiedu_reset_stats()
    L.fwiedu_set_kernel_timing(ktime_every)
    times = []
    misfit = None
    for i in range(steps):
        t0 = time.perf_counter()
        misfit = L.fwiedu_session_gradient(sess, None)  # returns after a stream sync
        times.append(time.perf_counter() - t0)
        log(f"step {i + 1}/{steps}: {times[-1] * 1e3:.1f} ms, misfit {misfit:.6g}")
    L.fwiedu_set_kernel_timing(0)
    names = ["fd_launches", "fd_inplace", "fd_timed", "fd_ms_avg", "fd_cells_per_launch",
             "fd_shots_per_launch", "batch", "device_bytes", "tile_tz", "tile_spb", "tile_kind",
             "tile_zfast", "tile_remap", "tile_txm", "tile_pd", "tile_wpb", "tile_zrun", "layout"]
    names += [f"fd{m}_{k}" for m in range(3) for k in
              ("timed", "ms_avg", "cells_per_launch", "shot_cells_per_launch",
               "acc_shot_cells_per_launch")]
    names += [f"tile{m}_{k}" for m in range(3) for k in
              ("kind", "tz", "spb", "txm", "zfast", "pd", "remap", "wpb", "zrun")]
    names += [f"tile{m}{v}_{k}" for m in (0, 2) for v in (1, 2) for k in ("zrun", "pd")]
    stats = {k: flib.stat(k) for k in names}
    L.fwiedu_session_destroy(sess)
    L.fwiedu_set_batch(0)
    return times, stats, misfit


# HBM bytes per launch of each stencil mode of the c1 gradient at 120 shots in
# lanes (forward and adjoint: averaged over the gradient's launches, 3200 plain,
# 700 staging and 101 adding launches of 4001 with steps in place; 600 and 201
# with three rotating fields, measured when three terms were staged per pass;
# the reconstruction's over its plain launches and those that store an imaging
# step's field a second time),
# per stencil shape (kernel kind 3: lanes with LDS row exchange; kind 4: the
# same with the register ring; columns per block, rows per wave, XCD order,
# block order) and time steps in place (1) or on three rotating fields (0);
# the rows of loads in flight change when a row's loads are issued, not which
# bytes are loaded: FETCH_SIZE x 2 + WRITE_SIZE from rocprofv3 counter runs
# scaled by a 512 MiB copy; the runs and the calibration are in
# profiles/fd_lanes_c1_traffic.md (kind 3) and profiles/fd_ring_c1_traffic.md
# (kind 4).  Kind 4's rotating-field forward and adjoint figures other than 20
# rows, XCD order, z-runs fastest were measured while the staging launches ran
# the LDS kernel (that shape read 0.4-0.5 % less then than with the staging
# in the ring stencil).
C1_LANES_TRAFFIC = {
    "forward": {(4, 8, 20, 1, 2, 1): 349.0e6,
                (3, 8, 16, 1, 0, 0): 353.8e6, (3, 8, 16, 1, 2, 0): 349.9e6, (3, 8, 16, 0, 0, 0): 443.6e6,
                (3, 8, 16, 0, 2, 0): 443.7e6, (3, 8, 20, 1, 0, 0): 360.0e6, (3, 8, 20, 1, 2, 0): 348.2e6,
                (3, 8, 20, 0, 0, 0): 431.6e6, (3, 8, 20, 0, 2, 0): 431.6e6, (3, 8, 24, 1, 0, 0): 357.4e6,
                (3, 8, 24, 1, 2, 0): 347.8e6, (3, 8, 24, 0, 0, 0): 426.0e6, (3, 8, 24, 0, 2, 0): 425.9e6,
                (3, 8, 32, 1, 0, 0): 351.0e6, (3, 8, 32, 1, 2, 0): 347.6e6, (3, 8, 32, 0, 0, 0): 420.1e6,
                (3, 8, 32, 0, 2, 0): 420.2e6, (4, 8, 20, 1, 0, 0): 364.5e6, (4, 8, 20, 1, 2, 0): 347.8e6,
                (4, 8, 20, 0, 0, 0): 445.0e6, (4, 8, 20, 0, 2, 0): 444.8e6, (4, 8, 30, 1, 0, 0): 359.0e6,
                (4, 8, 30, 1, 2, 0): 346.9e6, (4, 8, 30, 0, 0, 0): 429.5e6, (4, 8, 30, 0, 2, 0): 429.5e6},
    "reconstruction": {(4, 8, 20, 1, 2, 1): 330.1e6,
                       (3, 8, 16, 1, 0, 0): 318.1e6, (3, 8, 16, 1, 2, 0): 311.4e6,
                       (3, 8, 16, 0, 0, 0): 404.1e6, (3, 8, 16, 0, 2, 0): 404.2e6,
                       (3, 8, 20, 1, 0, 0): 323.9e6, (3, 8, 20, 1, 2, 0): 310.7e6,
                       (3, 8, 20, 0, 0, 0): 392.9e6, (3, 8, 20, 0, 2, 0): 393.0e6,
                       (3, 8, 24, 1, 0, 0): 320.6e6, (3, 8, 24, 1, 2, 0): 310.8e6,
                       (3, 8, 24, 0, 0, 0): 387.6e6, (3, 8, 24, 0, 2, 0): 387.7e6,
                       (3, 8, 32, 1, 0, 0): 313.4e6, (3, 8, 32, 1, 2, 0): 311.3e6,
                       (3, 8, 32, 0, 0, 0): 382.1e6, (3, 8, 32, 0, 2, 0): 382.1e6,
                       (4, 8, 20, 1, 0, 0): 329.1e6, (4, 8, 20, 1, 2, 0): 312.4e6,
                       (4, 8, 20, 0, 0, 0): 412.1e6, (4, 8, 20, 0, 2, 0): 412.0e6,
                       (4, 8, 30, 1, 0, 0): 324.7e6, (4, 8, 30, 1, 2, 0): 311.6e6,
                       (4, 8, 30, 0, 0, 0): 395.4e6, (4, 8, 30, 0, 2, 0): 395.4e6},
    "adjoint": {(4, 8, 20, 1, 2, 1): 366.2e6,
                (3, 8, 16, 1, 0, 0): 363.6e6, (3, 8, 16, 1, 2, 0): 370.6e6, (3, 8, 16, 0, 0, 0): 463.8e6,
                (3, 8, 16, 0, 2, 0): 463.9e6, (3, 8, 20, 1, 0, 0): 374.1e6, (3, 8, 20, 1, 2, 0): 366.5e6,
                (3, 8, 20, 0, 0, 0): 451.5e6, (3, 8, 20, 0, 2, 0): 451.7e6, (3, 8, 24, 1, 0, 0): 375.7e6,
                (3, 8, 24, 1, 2, 0): 365.6e6, (3, 8, 24, 0, 0, 0): 446.0e6, (3, 8, 24, 0, 2, 0): 445.8e6,
                (3, 8, 32, 1, 0, 0): 371.9e6, (3, 8, 32, 1, 2, 0): 364.2e6, (3, 8, 32, 0, 0, 0): 439.5e6,
                (3, 8, 32, 0, 2, 0): 439.7e6, (4, 8, 20, 1, 0, 0): 385.4e6, (4, 8, 20, 1, 2, 0): 368.0e6,
                (4, 8, 20, 0, 0, 0): 464.9e6, (4, 8, 20, 0, 2, 0): 464.7e6, (4, 8, 30, 1, 0, 0): 378.6e6,
                (4, 8, 30, 1, 2, 0): 366.7e6, (4, 8, 30, 0, 0, 0): 449.1e6, (4, 8, 30, 0, 2, 0): 449.0e6},
}


def measured_traffic(w, st, mode):
    m = MODE_INDEX[mode]
    kind = int(st[f"tile{m}_kind"])
    if w.name != "c1" or w.nshot != 120 or kind not in (3, 4):
        return None
    key = (kind, int(st[f"tile{m}_wpb"]), int(st[f"tile{m}_zrun"]), int(st[f"tile{m}_remap"]),
           int(st[f"tile{m}_zfast"]), 1 if st.get("fd_inplace", 0) > 0 else 0)
    return C1_LANES_TRAFFIC.get(mode, {}).get(key)


MODE_INDEX = {"forward": 0, "reconstruction": 1, "adjoint": 2}
MODE_TAGS = {"forward": "FD_FWD", "reconstruction": "FD_RECON", "adjoint": "FD_ADJ"}
# algorithmic bytes of a launch's accumulation epilogue per shot-cell: the
# pseudo-Hessian (FP64 read + write) of the forward, the imaging condition
# (forward field read, FP64 gradient read + write) of the adjoint
EPILOGUE_BYTES = {"forward": 16.0, "reconstruction": 0.0, "adjoint": 20.0}


def stencil_modes(w, st):
    """Per stencil mode, from the HIP-event samples of the timed gradients:
    average launch time and algorithmic bytes per launch.  The stencil moves
    12 B per shot-cell (p1, p0 read, p2 written) over the interior cells plus
    the 7 coefficient streams of the reference formulation, 28 B per cell,
    shared by the launch's shots; the forward and adjoint add their
    accumulation epilogue over the nx x nz model cells on the launches where it
    runs (every 5th step)."""
    out = {}
    total = sum(st[f"fd{m}_ms_avg"] * st[f"fd{m}_timed"] for m in range(3)) or 1.0
    for m, name in enumerate(("forward", "reconstruction", "adjoint")):
        n = int(st[f"fd{m}_timed"])
        ms = st[f"fd{m}_ms_avg"]
        cells = st[f"fd{m}_cells_per_launch"]
        model_share = (w.nx * w.nz) / cells if cells > 0 else 0.0
        b = (28.0 * cells + 12.0 * st[f"fd{m}_shot_cells_per_launch"] +
             EPILOGUE_BYTES[name] * st[f"fd{m}_acc_shot_cells_per_launch"] * model_share)
        ach = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[name] = {"avg_launch_us": ms * 1e3, "bytes_per_launch": b, "achieved": ach,
                     "frac": ach / HBM_PEAK_GBS if ach else None, "launches_timed": n,
                     "time_share": ms * n / total}
    return out


KERNELS = {0: "fd_kernel (square tiles, per-shot arenas)",
           1: "fd_march_kernel (column march, per-shot arenas)",
           2: "fd_lanes_kernel (shots in lanes)",
           3: "fd_lanes_lds_kernel (shots in lanes, LDS row exchange)",
           4: "fd_ring_kernel (shots in lanes, LDS row exchange, register ring)"}


def tile_shape(st, m):
    k = int(st[f"tile{m}_kind"])
    d = {"kernel": KERNELS[k], "xcd_order": int(st[f"tile{m}_remap"])}
    if k == 0:
        d.update(rows=int(st[f"tile{m}_tz"]), shots_per_block=int(st[f"tile{m}_spb"]),
                 z_fast_order=int(st[f"tile{m}_zfast"]))
    elif k == 1:
        d.update(cols=int(st[f"tile{m}_txm"]), shots_per_block=int(st[f"tile{m}_spb"]),
                 prefetch=int(st[f"tile{m}_pd"]))
    else:
        d.update(cols_per_block=int(st[f"tile{m}_wpb"]), rows_per_wave=int(st[f"tile{m}_zrun"]),
                 prefetch_rows=int(st[f"tile{m}_pd"]), block_order=int(st[f"tile{m}_zfast"]))
    return d


def stencil_shape(st):
    """The autotuned launch shape of each stencil mode of the last batch, and
    (forward, adjoint) the rows per wave and rows in flight of its launches that
    stage their accumulation term or add the staged terms to the accumulator."""
    d = {"layout": "shots in lanes" if st["layout"] else "per-shot arenas"}
    for name, m in MODE_INDEX.items():
        d[name] = tile_shape(st, m)
        if m != 1 and st.get(f"tile{m}_kind") == 3:
            for v, tag in ((1, "staging"), (2, "adding")):
                d[name][tag] = {"rows_per_wave": int(st[f"tile{m}{v}_zrun"]),
                                "prefetch_rows": int(st[f"tile{m}{v}_pd"])}
    return d


def copy_calibration(mb):
    """A plain device-to-device copy of known size, so that hardware byte
    counters (FETCH_SIZE, WRITE_SIZE) of the stencil can be scaled."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {mb} MiB device-to-device")


def stream_leg():
    """The GPU's HBM streaming rate for the stencil's traffic pattern (two
    float32 streams read, one written, each element once), measured by
    fwi-edu_amd/bin/hbm_stream in a child process at array sizes inside and
    beyond the 256 MB Infinity Cache: the rate a plain streaming kernel reaches
    here, beside which the stencil's fraction of the 8 TB/s peak reads."""
    exe = os.path.join(ROOT, "fwi-edu_amd", "bin", "hbm_stream")
    if not os.path.exists(exe):
        log("hbm_stream probe not built (make -C fwi-edu_amd): stream rates omitted")
        return None
    r = subprocess.run([exe, "3"], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"hbm_stream probe failed ({r.returncode}): {r.stderr.strip()[-300:]}")
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    beyond = [z for z in res["sizes"] if z["array_mib"] >= 700]
    res["add_beyond_cache_tb_s"] = max(z["add_tb_s"] for z in beyond)
    log(f"hbm stream: add {res['add_beyond_cache_tb_s']:.2f} TB/s beyond the Infinity Cache")
    return res


def serving_leg(args):
    """Continuous-batching serving loop: a seeded Poisson stream of modeling /
    gradient requests (70/30) on random-layer models, grids uniform in 128-1024
    per dimension at 10 m, nt uniform in 1000-4000, 1-64 shots, submitted at
    their arrival times to a server with `--serving-slots` workers."""
    from fwiedu import serving
    stream = serving.request_stream(args.serving, args.serving_rate, 303)
    log(f"serving: {args.serving} requests at {args.serving_rate}/s "
        f"(each built shortly before its arrival)")
    srv = serving.Server(slots=args.serving_slots, mem_fraction=0.8, graphs=True)
    try:
        jobs, wall = serving.run_stream(srv, stream, log=log)
        out = {"requests": args.serving, "rate_per_s": args.serving_rate,
               "slots": args.serving_slots,
               "gradient_jobs": sum(1 for j in jobs if j.problem == 1),
               "shots": sum(j.shots() for j in jobs), "wall_s": wall,
               "shots_per_s": sum(j.shots() for j in jobs) / wall,
               "jobs_per_s": len(jobs) / wall,
               "p50_latency_s": srv.stat("p50_s"), "p99_latency_s": srv.stat("p99_s"),
               "latency_ranks": f"nearest rank over these {args.serving} requests: p99 is "
                                f"latency number {int(np.ceil(0.99 * args.serving))} of "
                                f"{args.serving} in increasing order (the full 2000-request "
                                f"stream: --serving 2000)",
               "max_latency_s": srv.stat("max_s"),
               "mean_queue_s": float(np.mean([j.queue for j in jobs])),
               "peak_concurrent_jobs": int(srv.stat("peak_concurrency")),
               "stream": "seed 303; grids 128-1024 x 128-1024 at 10 m, dt 0.8 ms, Ricker 10 Hz; "
                         "nt 1000-4000; 1-64 shots; 30 % gradient jobs (observed traces zero)"}
    finally:
        srv.close()
    if args.serving_log:
        with open(args.serving_log, "w") as f:
            for j in jobs:
                f.write(json.dumps(j.row()) + "\n")
    log(f"serving: p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s, "
        f"{out['shots_per_s']:.1f} shots/s")
    return out


def large_leg(L, args):
    """Stencil launches on a working set beyond the 256 MB Infinity Cache: the
    400 x 1200 multiscale workload (nt 5000 cut to --large-nt steps: a launch's
    work and bytes do not depend on nt; 1500 steps, 1.2 s, let the sea-floor
    reflections reach the receivers, so the residual and the adjoint fields
    are not zero), 256 shots in one batch, one warm-up gradient (launch-shape
    tuning) and one timed gradient.  Evidence for the stencil's HBM fraction
    at this size; not the bench value."""
    w = synth.workload("c2", nt=args.large_nt)
    log(f"large leg {w.name}: {w.nz}x{w.nx} nt={w.nt} shots={w.nshot}")
    obs = observed(L, w)
    times, st, misfit = timed_gradients(L, w, obs, 1, 1, 0, 7)
    modes = stencil_modes(w, st)
    for name, m in modes.items():
        log(f"large leg {name}: {m['avg_launch_us']:.1f} us, {m['frac']:.3f} of HBM peak")
    dom = max(modes, key=lambda m: modes[m]["time_share"])
    return {"workload": f"{w.name}: {w.nz}x{w.nx} @ {w.dx} m, {w.nshot} shots, free surface "
                        f"{w.freesurface}, gradient of nt {w.nt} (a launch's bytes and time do "
                        f"not depend on nt)",
            "misfit": misfit, "gradient_shots_per_s": w.nshot / times[0],
            "working_set_bytes_per_launch": modes[dom]["bytes_per_launch"],
            "shots_per_launch": int(st["fd_shots_per_launch"]),
            "stencil": stencil_shape(st), "stencil_modes": modes,
            "roofline": {"bound": "hbm", "achieved": modes[dom]["achieved"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": modes[dom]["frac"], "mode": dom}}


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family / model /
    stepping numbers and the logical CPU count."""
    info = {}
    n = 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "processor":
                    n += 1
                elif k in ("model name", "vendor_id", "cpu family", "model", "stepping",
                           "cpu MHz", "cache size") and k not in info:
                    info[k] = v.strip()
    except OSError:
        return "unknown"
    return (f"{info.get('model name', 'unknown')} ({info.get('vendor_id', '?')} family "
            f"{info.get('cpu family', '?')} model {info.get('model', '?')} stepping "
            f"{info.get('stepping', '?')}, {n} logical CPUs, {info.get('cache size', '?')} cache)")


def native_oracle():
    """The FP32 oracle built on this host with the reference's flags
    (-O3 -march=native -ffast-math -fopenmp, Makefile.gcc:28); None (and the
    prebuilt -march=x86-64-v3 build is used) if the build fails."""
    import tempfile
    d = tempfile.mkdtemp(prefix="fwiedu_oracle_")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                        f"NATIVE_DIR={d}"], check=True, capture_output=True, timeout=120)
        return os.path.join(d, "liboracle_f32n.so")
    except (subprocess.SubprocessError, OSError):
        return None


def cpu_baseline(w, nshots, shots_1t):
    """The FP32 oracle (reference algorithm and compiler flags, OpenMP over all
    host threads) on `nshots` gradient shots of the same workload, then on
    `shots_1t` of them with one OpenMP thread."""
    # thread placement as the baseline plan states; libgomp reads these when
    # the oracle library first loads it
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import Oracle
    native = native_oracle()
    o = Oracle("f32", path=native) if native else Oracle("f32")
    build = ("-O3 -march=native -ffast-math, built on this host" if native else
             "-O3 -march=x86-64-v3 -ffast-math (prebuilt; the native build failed)")
    sub = synth.workload(w.name, nshot=nshots)
    obs = sub.shots()
    vpt, rht = Array2D(sub.vp_true), Array2D(sub.rho_true)
    o.propagation(vpt.pp, rht.pp, obs.recs, nshots, sub.nz, sub.nx, sub.dx, sub.dz, sub.dt, sub.nt,
                  sub.pml, sub.freesurface, 0.0, 2.0, 1, 0, obs.recs, None)
    vps, rhs = Array2D(sub.vp_start), Array2D(sub.rho_start)

    def gradient(n):
        pre = obs.header_copy()
        g = Array2D(np.zeros((sub.nx, sub.nz), np.float32))
        t0 = time.perf_counter()
        o.propagation(vps.pp, rhs.pp, obs.recs, n, sub.nz, sub.nx, sub.dx, sub.dz, sub.dt, sub.nt,
                      sub.pml, sub.freesurface, 0.0, 2.0, 1, 1, pre.recs, g.pp)
        return time.perf_counter() - t0

    dt = gradient(nshots)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": nshots / dt, "unit": "gradient shots/s", "cores": cores, "kind": "port",
           "sample": f"{nshots} of the {w.nshot} gradient shots of the same workload "
                     f"({dt:.1f} s), the FP32 oracle ({build}, OpenMP, "
                     f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                     f"OMP_PLACES={os.environ.get('OMP_PLACES')})",
           "cpu": cpu_model()}
    if shots_1t > 0:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        try:
            dt1 = gradient(min(shots_1t, nshots))
        finally:
            gomp.omp_set_num_threads(cores)
        out["value_1_thread"] = min(shots_1t, nshots) / dt1
        out["sample_1_thread"] = f"{min(shots_1t, nshots)} gradient shot(s), one OpenMP thread ({dt1:.1f} s)"
    log(f"cpu baseline: {out['value']:.3f} shots/s on {cores} threads"
        + (f", {out['value_1_thread']:.4f} on 1" if "value_1_thread" in out else ""))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--nshot", type=int, default=0, help="override the shot count")
    ap.add_argument("--nt", type=int, default=0,
                    help="override the time steps (short profiled runs; not a bench line)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this size (counter calibration)")
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0 = all that fit)")
    ap.add_argument("--plain-shots", type=int, default=4)
    ap.add_argument("--graph-steps", type=int, default=2,
                    help="gradients timed with the time loops replayed from HIP graphs (0 = skip)")
    ap.add_argument("--cpu-shots", type=int, default=4)
    ap.add_argument("--cpu-1thread-shots", type=int, default=1,
                    help="gradient shots of the CPU baseline timed on one thread (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the HBM streaming-rate probe")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-array (PCIe-inclusive) iso_acoustic2d_propagation timing")
    ap.add_argument("--ktime-every", type=int, default=41,
                    help="HIP-event timing of every Nth stencil launch of each mode (a timed "
                         "launch leaves a few microseconds of idle device time around it; "
                         "41 is prime to the accumulation period of 5)")
    ap.add_argument("--serving", type=int, default=48,
                    help="requests of the serving leg (0 = skip; 2000 = the full stream)")
    ap.add_argument("--serving-rate", type=float, default=3.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--serving-log", default="",
                    help="write one JSON line per completed request (sizes, latency, queueing)")
    ap.add_argument("--ckpt-bits", type=int, default=32, choices=(32, 16),
                    help="checkpoint storage: FP32 or FP16 with a per-snapshot scale")
    ap.add_argument("--freq", type=float, default=0.0,
                    help="band frequency of the adjoint-source filter (0 = no filter)")
    ap.add_argument("--filter", type=int, default=0, choices=(0, 1),
                    help="with --freq: 0 low-pass, 1 band-pass")
    ap.add_argument("--large-nt", type=int, default=1500,
                    help="time steps of the 400x1200, 256-shot stencil leg beyond the Infinity "
                         "Cache (0 = skip)")
    ap.add_argument("--lanes", default="",
                    help="force the shots-in-lanes stencil shape WPB,ZRUN,PD,XCD[,LDS[,ORDER]] "
                         "(LDS 0 none, 1 LDS row exchange, 2 register ring; ORDER the block "
                         "order 0, 1 or 2; default: autotuned)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (the shots "
              "of a gradient would be split across GPUs, one process each, with one FP64 "
              "all-reduce of gradient and pseudo-Hessian per gradient)", file=sys.stderr)
        sys.exit(2)
    L = flib.load()
    if not L.fwiedu_device_ok():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    L.fwiedu_set_checkpoint_bits(args.ckpt_bits)
    if args.lanes:
        v = [int(x) for x in args.lanes.split(",")]
        wpb, zrun, pd, xcd = v[:4]
        L.fwiedu_set_xcd_remap(xcd)
        L.fwiedu_set_tile_order(v[5] if len(v) > 5 else 0)
        L.fwiedu_set_lanes(wpb, zrun, pd, v[4] if len(v) > 4 else 0)
    over = {"nshot": args.nshot} if args.nshot else {}
    if args.nt:
        over["nt"] = args.nt
    if args.freq > 0:
        # one band of the multiscale sequence: FFT filter of the adjoint sources
        over["freq"] = args.freq
        over["filter"] = args.filter
    w = synth.workload(args.workload, **over)
    log(f"workload {w.name}: {w.nz}x{w.nx} nt={w.nt} shots={w.nshot}")
    t0 = time.perf_counter()
    obs = observed(L, w)
    log(f"observed data modelled on the GPU in {time.perf_counter() - t0:.2f} s")

    times, st, misfit = timed_gradients(L, w, obs, args.steps, args.warmup, args.batch,
                                        args.ktime_every)
    step_s = float(np.mean(times))
    shots_per_s = w.nshot / step_s
    cu = w.cell_updates_per_gradient_shot()
    S = int(st["fd_shots_per_launch"])
    modes = stencil_modes(w, st)
    dom = max(modes, key=lambda m: modes[m]["time_share"])

    # the reference boundary as the FWI driver calls it: host models and host
    # records in, host adjoint sources and host gradient out (PCIe included)
    host = None
    if not args.no_host:
        vps, rhs = Array2D(w.vp_start), Array2D(w.rho_start)
        pre = obs.header_copy()
        g = Array2D(np.zeros((w.nx, w.nz), np.float32))
        t0 = time.perf_counter()
        flib.propagate(L, w, vps, rhs, obs, 1, pre, g)
        th = time.perf_counter() - t0
        host = {"gradient_shots_per_s": w.nshot / th, "ms": th * 1e3,
                "call": "iso_acoustic2d_propagation(problem=1) on host arrays, 1 call"}
        log(f"host-boundary gradient: {w.nshot / th:.2f} shots/s")

    # the same gradients with each batch's time loops captured into HIP graphs
    # and replayed (one hipGraphLaunch per loop; sampled kernel timing skips
    # replayed launches, so the stencil figures above come from the eager runs)
    graph = None
    if args.graph_steps > 0:
        L.fwiedu_set_graphs(1)
        try:
            # two warm-up gradients: the first runs a loop as it is, the
            # second captures it; the timed ones replay it
            gtimes, _, _ = timed_gradients(L, w, obs, args.graph_steps, 2, args.batch, 0)
            launches = flib.stat("graph_launches")
        finally:
            L.fwiedu_set_graphs(0)
        gs = w.nshot / float(np.mean(gtimes))
        graph = {"gradient_shots_per_s": gs, "ms_per_step": float(np.mean(gtimes)) * 1e3,
                 "graph_launches": int(launches),
                 "sample": f"{args.graph_steps} timed gradients after 2 warm-up (eager, capture)"}
        log(f"graph replay: {gs:.2f} shots/s")

    # plain path: one shot per launch
    plain = None
    if args.plain_shots > 0:
        pw = synth.workload(args.workload, nshot=args.plain_shots)
        pobs = observed(L, pw)
        # one warm-up gradient first: launch-shape tuning of the one-shot
        # geometry stays out of the timed one
        ptimes, pst, _ = timed_gradients(L, pw, pobs, 1, 1, 1, args.ktime_every)
        ps = pw.nshot / float(np.mean(ptimes))
        plain = {"shots_per_launch": 1, "gradient_shots_per_s": ps,
                 "gcell_updates_per_s": ps * cu / 1e9,
                 "fd_avg_launch_us": pst["fd_ms_avg"] * 1e3,
                 "sample": f"{args.plain_shots} shots, 1 warm-up and 1 timed gradient"}
        log(f"plain: {ps:.2f} shots/s")

    serve = serving_leg(args) if args.serving > 0 else None
    large = large_leg(L, args) if args.large_nt > 0 and not args.lanes else None

    if args.copy_mb:
        copy_calibration(args.copy_mb)
    stream = None if args.no_stream else stream_leg()
    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_shots, args.cpu_1thread_shots)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3f} shots/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": shots_per_s,
        "unit": "gradient shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model with faults; observed data modelled on the GPU)",
        "config": {"workload": f"{w.name}: {w.nz}x{w.nx} @ {w.dx} m, nt {w.nt}, {w.nshot} shots x "
                               f"{len(w.receiver_columns())} receivers, pml {w.pml}, free surface "
                               f"{w.freesurface}, checkpointed adjoint gradient, precondition "
                               f"{w.precondition}, model and records resident in HBM" +
                               (f", adjoint sources {'low' if w.filter == 0 else 'band'}-pass "
                                f"filtered at {w.freq:g} Hz" if w.freq > 0 else ""),
                   "shots_per_launch": S, "batch": int(st["batch"]),
                   "checkpoint_storage": f"fp{args.ckpt_bits}",
                   "stencil": stencil_shape(st)},
        "gcell_updates_per_s": shots_per_s * cu / 1e9,
        "cell_updates_per_gradient_shot": cu,
        "misfit": misfit,
        "plain": plain,
        "batched_over_plain": (shots_per_s / plain["gradient_shots_per_s"]) if plain else None,
        "host_boundary": host,
        "graph_replay": graph,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": modes[dom]["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": modes[dom]["frac"],
                     "traffic": measured_traffic(w, st, dom),
                     "traffic_source": "profiles/fd_lanes_c1_traffic.md, "
                                       "profiles/fd_ring_c1_traffic.md (counter runs of the same "
                                       "stencil shape; null when the autotuner picked a shape not "
                                       "measured there)",
                     "kernel": KERNELS[int(st[f"tile{MODE_INDEX[dom]}_kind"])].split(" ")[0] +
                               f"<{MODE_TAGS[dom]}> "
                               f"({dom}: the stencil mode with the largest share of kernel time)",
                     "bytes_per_launch": modes[dom]["bytes_per_launch"],
                     "avg_launch_us": modes[dom]["avg_launch_us"],
                     "launches_timed": modes[dom]["launches_timed"],
                     # the same achieved rate over what a plain streaming
                     # kernel of the stencil's pattern reaches beyond the
                     # Infinity Cache on this GPU (hbm_stream below)
                     "frac_of_stream": (modes[dom]["achieved"] /
                                        (stream["add_beyond_cache_tb_s"] * 1e3)) if stream else None},
        "stencil_modes": modes,
        "beyond_infinity_cache": large,
        "hbm_stream": stream,
        "cpu_baseline": cpu,
        "device_bytes": st["device_bytes"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
