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
import json
import os
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
    stats = {k: flib.stat(k) for k in ("fd_launches", "fd_timed", "fd_ms_avg", "fd_cells_per_launch",
                                       "fd_shots_per_launch", "batch", "device_bytes", "tile_tz",
                                       "tile_spb", "tile_kind", "tile_zfast", "tile_remap",
                                       "tile_txm", "tile_pd", "tile_wpb", "tile_zrun", "layout")}
    L.fwiedu_session_destroy(sess)
    L.fwiedu_set_batch(0)
    return times, stats, misfit


# HBM bytes of one reconstruction launch of the c1 workload at 120 shots in
# lanes, per stencil shape (columns per block, rows per wave, rows in flight,
# XCD order): FETCH_SIZE x 2 + WRITE_SIZE from rocprofv3 counter runs scaled by
# a 512 MiB copy; the runs and the calibration are in
# profiles/fd_lanes_recon_c1_traffic.md.
C1_LANES_TRAFFIC = {(2, 8, 8, 1, 0): 447.5e6, (2, 8, 16, 1, 1): 345.5e6, (2, 4, 8, 1, 0): 518.8e6,
                    (2, 8, 8, 2, 0): 447.3e6, (3, 8, 16, 2, 0): 413.9e6, (3, 8, 16, 2, 1): 352.4e6}


def measured_traffic(w, st):
    kind = int(st["tile_kind"])
    if w.name != "c1" or w.nshot != 120 or kind not in (2, 3) or int(st["tile_zfast"]) != 0:
        return None
    key = (kind, int(st["tile_wpb"]), int(st["tile_zrun"]), int(st["tile_pd"]),
           int(st["tile_remap"]))
    return C1_LANES_TRAFFIC.get(key)


KERNELS = {0: "fd_kernel (square tiles, per-shot arenas)",
           1: "fd_march_kernel (column march, per-shot arenas)",
           2: "fd_lanes_kernel (shots in lanes)",
           3: "fd_lanes_lds_kernel (shots in lanes, LDS row exchange)"}


def stencil_shape(st):
    k = int(st["tile_kind"])
    d = {"kernel": KERNELS[k], "layout": "shots in lanes" if st["layout"] else "per-shot arenas",
         "xcd_order": int(st["tile_remap"])}
    if k == 0:
        d.update(rows=int(st["tile_tz"]), shots_per_block=int(st["tile_spb"]),
                 z_fast_order=int(st["tile_zfast"]))
    elif k == 1:
        d.update(cols=int(st["tile_txm"]), shots_per_block=int(st["tile_spb"]),
                 prefetch=int(st["tile_pd"]))
    else:
        d.update(cols_per_block=int(st["tile_wpb"]), rows_per_wave=int(st["tile_zrun"]),
                 prefetch_rows=int(st["tile_pd"]), block_order=int(st["tile_zfast"]))
    return d


def copy_calibration(mb):
    """A plain device-to-device copy of known size, so that hardware byte
    counters (FETCH_SIZE, WRITE_SIZE) of the stencil can be scaled."""
    import ctypes
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


def serving_leg(args):
    """Continuous-batching serving loop: a seeded Poisson stream of modeling /
    gradient requests (70/30) on random-layer models, grids uniform in 128-1024
    per dimension at 10 m, nt uniform in 1000-4000, 1-64 shots, submitted at
    their arrival times to a server with `--serving-slots` workers."""
    from fwiedu import serving
    t0 = time.perf_counter()
    stream = serving.request_stream(args.serving, args.serving_rate, 303)
    log(f"serving: {args.serving} requests built in {time.perf_counter() - t0:.1f} s")
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
               "max_latency_s": srv.stat("max_s"),
               "mean_queue_s": float(np.mean([j.queue for j in jobs])),
               "peak_concurrent_jobs": int(srv.stat("peak_concurrency")),
               "stream": "seed 303; grids 128-1024 x 128-1024 at 10 m, dt 0.8 ms, Ricker 10 Hz; "
                         "nt 1000-4000; 1-64 shots; 30 % gradient jobs"}
    finally:
        srv.close()
    log(f"serving: p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s, "
        f"{out['shots_per_s']:.1f} shots/s")
    return out


def cpu_baseline(w, nshots):
    """The FP32 oracle (reference algorithm and compiler flags, OpenMP over all
    host threads) on `nshots` gradient shots of the same workload."""
    from oracle import Oracle
    o = Oracle("f32")
    sub = synth.workload(w.name, nshot=nshots)
    obs = sub.shots()
    vpt, rht = Array2D(sub.vp_true), Array2D(sub.rho_true)
    o.propagation(vpt.pp, rht.pp, obs.recs, nshots, sub.nz, sub.nx, sub.dx, sub.dz, sub.dt, sub.nt,
                  sub.pml, sub.freesurface, 0.0, 2.0, 1, 0, obs.recs, None)
    vps, rhs = Array2D(sub.vp_start), Array2D(sub.rho_start)
    pre = obs.header_copy()
    g = Array2D(np.zeros((sub.nx, sub.nz), np.float32))
    t0 = time.perf_counter()
    o.propagation(vps.pp, rhs.pp, obs.recs, nshots, sub.nz, sub.nx, sub.dx, sub.dz, sub.dt, sub.nt,
                  sub.pml, sub.freesurface, 0.0, 2.0, 1, 1, pre.recs, g.pp)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nshots / dt, "unit": "gradient shots/s", "cores": cores, "kind": "port",
            "sample": f"{nshots} of the 120 gradient shots of the same workload "
                      f"({dt:.1f} s), oracle/liboracle_f32.so (-O3 -ffast-math, OpenMP)"}


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
    ap.add_argument("--cpu-shots", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-array (PCIe-inclusive) iso_acoustic2d_propagation timing")
    ap.add_argument("--ktime-every", type=int, default=11)
    ap.add_argument("--serving", type=int, default=48,
                    help="requests of the serving leg (0 = skip; 2000 = the full stream)")
    ap.add_argument("--serving-rate", type=float, default=3.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--ckpt-bits", type=int, default=32, choices=(32, 16),
                    help="checkpoint storage: FP32 or FP16 with a per-snapshot scale")
    ap.add_argument("--lanes", default="",
                    help="force the shots-in-lanes stencil shape WPB,ZRUN,PD,XCD[,LDS] "
                         "(default: autotuned)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots would shard across GPUs, see DESIGN.md)", file=sys.stderr)
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
        L.fwiedu_set_lanes(wpb, zrun, pd, v[4] if len(v) > 4 else 0)
    over = {"nshot": args.nshot} if args.nshot else {}
    if args.nt:
        over["nt"] = args.nt
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
    cells = interior_cells(w)
    # algorithmic bytes of one launch: p1, p0 read and p2 written per shot-cell
    # (12 B), plus the 7 coefficient streams of the reference formulation
    # (28 B per cell) shared by the S shots of the launch
    bytes_per_launch = cells * (12.0 * S + 28.0)
    avg_ms = st["fd_ms_avg"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None

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

    # plain path: one shot per launch
    plain = None
    if args.plain_shots > 0:
        pw = synth.workload(args.workload, nshot=args.plain_shots)
        pobs = observed(L, pw)
        ptimes, pst, _ = timed_gradients(L, pw, pobs, 1, 0, 1, args.ktime_every)
        ps = pw.nshot / float(np.mean(ptimes))
        plain = {"shots_per_launch": 1, "gradient_shots_per_s": ps,
                 "gcell_updates_per_s": ps * cu / 1e9,
                 "fd_avg_launch_us": pst["fd_ms_avg"] * 1e3,
                 "sample": f"{args.plain_shots} shots, 1 step"}
        log(f"plain: {ps:.2f} shots/s")

    serve = serving_leg(args) if args.serving > 0 else None

    if args.copy_mb:
        copy_calibration(args.copy_mb)
    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_shots)
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
                               f"{w.precondition}, model and records resident in HBM",
                   "shots_per_launch": S, "batch": int(st["batch"]),
                   "checkpoint_storage": f"fp{args.ckpt_bits}",
                   "stencil": stencil_shape(st)},
        "gcell_updates_per_s": shots_per_s * cu / 1e9,
        "cell_updates_per_gradient_shot": cu,
        "misfit": misfit,
        "plain": plain,
        "batched_over_plain": (shots_per_s / plain["gradient_shots_per_s"]) if plain else None,
        "host_boundary": host,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(w, st),
                     "traffic_source": "profiles/fd_lanes_recon_c1_traffic.md (counter runs of the "
                                       "same stencil shape; null when the autotuner picked a shape "
                                       "not measured there)",
                     "kernel": KERNELS[int(st["tile_kind"])].split(" ")[0] +
                               "<FD_RECON> (reconstruction sweep)",
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_ms * 1e3,
                     "launches_timed": int(st["fd_timed"])},
        "cpu_baseline": cpu,
        "device_bytes": st["device_bytes"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
