"""Serving loop: jobs of uneven grid sizes submitted together run overlapped on
the device (several workers busy at once), and every job's records, adjoint
sources and gradient are bit-identical to the same job run alone through
iso_acoustic2d_propagation (the stencil and the FP64 sums do not depend on which
other jobs share the device, on the batch size the admission left a job, or on
the launch shape its worker tuned); the misfit agrees to 1e-6 relative (the
server sums per shot, calculate_objective_value over all samples at once).
"""
import sys

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu import serving
from fwiedu.records import Array2D

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    lib = flib.load()
    assert lib.fwiedu_device_ok() == 1, "no HIP device"
    return lib


# (problem, nz, nx, nt, shots): uneven grids, both problems, 1-64 shots (the
# 50- and 64-shot jobs run shots in lanes, their time steps in place)
MIX = [(0, 60, 90, 400, 3), (1, 48, 70, 300, 4), (0, 100, 64, 500, 12), (1, 72, 120, 350, 2),
       (0, 40, 40, 300, 1), (1, 90, 60, 420, 5), (0, 128, 150, 300, 6), (1, 64, 64, 260, 1),
       (1, 56, 80, 300, 50), (0, 64, 96, 280, 64)]


def alone(L, job):
    """The same job through the reference entry point on the default context."""
    w = job.w
    pre = job.obs.header_copy()
    g = Array2D(np.zeros((w.nx, w.nz), np.float32)) if job.problem == 1 else None
    flib.propagate(L, w, job.vp, job.rho, job.obs, job.problem, pre, g)
    mis = L.calculate_objective_value(pre.recs, w.nshot) if job.problem == 1 else None
    return np.concatenate([t.ravel() for t in pre.all_traces()]), (g.a.copy() if g else None), mis


@pytest.mark.parametrize("bits", [32, 16])
def test_jobs_same_alone_and_in_the_mix(L, bits):
    say(f"case: serving mix of {len(MIX)} jobs, 3 slots, checkpoints fp{bits}")
    jobs = []
    for i, (p, nz, nx, nt, ns) in enumerate(MIX):
        w = serving.request_workload(i, 4242 + i, nz, nx, nt, ns)
        jobs.append(serving.Job(w, p, serving.observed_for(L, w) if p == 1 else None))
    srv = serving.Server(slots=3, mem_fraction=0.5, graphs=True, ckpt_bits=bits)
    try:
        for j in jobs:
            srv.submit(j)
        for j in jobs:
            srv.wait(j)
        peak = srv.stat("peak_concurrency")
        say(f"  peak concurrency {peak:.0f}, p50 {srv.stat('p50_s') * 1e3:.1f} ms, "
            f"p99 {srv.stat('p99_s') * 1e3:.1f} ms")
    finally:
        srv.close()
    assert peak >= 2, "jobs did not overlap"
    L.fwiedu_set_checkpoint_bits(bits)
    try:
        for j in jobs:
            L.fwiedu_reset_stats()
            rec, g, mis = alone(L, j)
            if j.w.nshot >= 48:
                # shots in lanes, the time steps in place
                assert flib.stat("layout") == 1 and flib.stat("fd_inplace") > 0, j.w.name
            assert np.array_equal(j.traces(), rec), j.w.name
            if j.problem == 1:
                assert np.array_equal(j.grad.a, g), j.w.name
                # the server sums per shot then over shots in FP64, the entry
                # point in one reduction: equal to FP32 resolution
                assert abs(j.misfit - mis) <= 1e-6 * abs(mis), (j.w.name, j.misfit, mis)
    finally:
        L.fwiedu_set_checkpoint_bits(32)


def test_jobs_same_with_scarce_memory(L):
    """A server whose device budget (90 MB) holds one job at a time and only
    some shots of the 5-shot gradient: jobs wait for memory (an idle worker
    gives its buffers back to a waiting job) and that gradient runs in
    batches of fewer shots; every result is still the job's own alone."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    frac = 90e6 / float(free.value)
    say(f"case: serving mix of {len(MIX)} jobs, 3 slots, 90 MB budget (fraction {frac:.2e})")
    jobs = []
    for i, (p, nz, nx, nt, ns) in enumerate(MIX):
        w = serving.request_workload(i, 5151 + i, nz, nx, nt, ns)
        jobs.append(serving.Job(w, p, serving.observed_for(L, w) if p == 1 else None))
    srv = serving.Server(slots=3, mem_fraction=frac, graphs=True)
    try:
        for j in jobs:
            srv.submit(j)
        for j in jobs:
            srv.wait(j)
        say(f"  peak concurrency {srv.stat('peak_concurrency'):.0f}, "
            f"mean queueing {np.mean([j.queue for j in jobs]) * 1e3:.1f} ms")
    finally:
        srv.close()
    for j in jobs:
        rec, g, mis = alone(L, j)
        assert np.array_equal(j.traces(), rec), j.w.name
        if j.problem == 1:
            assert np.array_equal(j.grad.a, g), j.w.name
