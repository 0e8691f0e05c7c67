"""Serving loop client: ctypes binding of the library's fwiedu_server_* entry
points and a seeded Poisson stream of modeling / gradient requests.

The server itself (admission by device footprint, one HIP stream per worker,
overlapping jobs) is native code in csrc/server.cpp; this module only builds
jobs, submits them at their arrival times and collects latencies.
"""
import ctypes
import threading
import time
from ctypes import POINTER, c_double, c_float, c_int, c_long, c_void_p

import numpy as np

from . import lib as flib
from . import synth
from .records import Array2D, Record2D, c_float_pp


class FwieduJob(ctypes.Structure):
    """FwieduJob of include/fwiedu.h."""
    _fields_ = [
        ("problem", c_int),
        ("vp", c_float_pp), ("rho", c_float_pp),
        ("obs", POINTER(Record2D)), ("pre", POINTER(Record2D)),
        ("grad", c_float_pp),
        ("nshot", c_int), ("nz", c_int), ("nx", c_int),
        ("dx", c_float), ("dz", c_float), ("dt", c_float),
        ("nt", c_int), ("pml", c_int), ("freesurface", c_int),
        ("freq", c_float),
        ("precondition", c_int),
    ]


def _bind(L):
    if getattr(L, "_server_bound", False):
        return L
    J = POINTER(FwieduJob)
    flib._sig(L, "fwiedu_server_create", c_void_p, [c_int, c_double, c_int, c_int])
    flib._sig(L, "fwiedu_server_submit", c_long, [c_void_p, J])
    flib._sig(L, "fwiedu_server_done", c_int, [c_void_p, c_long])
    flib._sig(L, "fwiedu_server_wait", c_int,
              [c_void_p, c_long, POINTER(c_double), POINTER(c_double), POINTER(c_double)])
    flib._sig(L, "fwiedu_server_stat", c_double, [c_void_p, ctypes.c_char_p])
    flib._sig(L, "fwiedu_server_destroy", None, [c_void_p])
    L._server_bound = True
    return L


class Job:
    """One request: a workload, its problem (0 modeling, 1 gradient) and the
    host arrays the server reads and writes (kept alive here)."""

    def __init__(self, w, problem, obs=None):
        self.w, self.problem = w, problem
        if problem == 1:
            assert obs is not None, "a gradient job needs observed records"
            self.vp, self.rho = Array2D(w.vp_start), Array2D(w.rho_start)
            self.obs = obs
        else:
            self.vp, self.rho = Array2D(w.vp_true), Array2D(w.rho_true)
            self.obs = w.shots()
        self.pre = self.obs.header_copy()
        self.grad = Array2D(np.zeros((w.nx, w.nz), np.float32)) if problem == 1 else None
        self.s = FwieduJob(problem, self.vp.pp, self.rho.pp, self.obs.recs, self.pre.recs,
                           self.grad.pp if self.grad is not None else None, w.nshot, w.nz, w.nx,
                           w.dx, w.dz, w.dt, w.nt, w.pml, w.freesurface, w.freq, w.precondition)
        self.id = None
        self.latency = self.queue = self.misfit = None

    def traces(self):
        return np.concatenate([t.ravel() for t in self.pre.all_traces()])

    def shots(self):
        return self.w.nshot


class Server:
    def __init__(self, slots=3, mem_fraction=0.8, graphs=True, ckpt_bits=32):
        self.L = _bind(flib.load())
        self.h = self.L.fwiedu_server_create(slots, mem_fraction, 1 if graphs else 0, ckpt_bits)

    def submit(self, job):
        job.id = self.L.fwiedu_server_submit(self.h, ctypes.byref(job.s))
        return job.id

    def wait(self, job):
        lat, q, mis = c_double(), c_double(), c_double()
        rc = self.L.fwiedu_server_wait(self.h, job.id, ctypes.byref(lat), ctypes.byref(q),
                                       ctypes.byref(mis))
        assert rc == 0
        job.latency, job.queue, job.misfit = lat.value, q.value, mis.value
        return job

    def stat(self, name):
        return self.L.fwiedu_server_stat(self.h, name.encode())

    def close(self):
        if self.h:
            self.L.fwiedu_server_destroy(self.h)
            self.h = None


def request_workload(i, seed, nz, nx, nt, nshot, dx=10.0):
    """A seeded random-layer request at dx = dz = 10 m, dt 0.8 ms, Ricker 10 Hz,
    sources and receivers one cell deep, MEAL 40, free surface on."""
    w = synth.Workload(name=f"req{i}", nz=nz, nx=nx, dx=dx, dz=dx, dt=0.0008, nt=nt, pml=40,
                       freesurface=1, nshot=nshot, shot_depth_cells=1, rcv_depth_cells=1,
                       f0=10.0, t0=0.15, seed=seed, waterdepth=max(4, nz // 12),
                       vpmin=1400.0, vpmax=5000.0, kind="layered")
    return w.build()


def request_stream(n, rate, seed, nz=(128, 1024), nx=(128, 1024), nt=(1000, 4000),
                   shots=(1, 64), p_gradient=0.3):
    """n requests with Poisson arrivals at `rate` per second: (arrival time,
    problem, workload parameters); sizes uniform in the given inclusive ranges.
    Workloads are built when a request is submitted (request_workload(*p))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        problem = 1 if rng.random() < p_gradient else 0
        p = (i, seed * 1000 + i, int(rng.integers(nz[0], nz[1] + 1)),
             int(rng.integers(nx[0], nx[1] + 1)), int(rng.integers(nt[0], nt[1] + 1)),
             int(rng.integers(shots[0], shots[1] + 1)))
        out.append((float(t[i] - t[0]), problem, p))
    return out


def observed_for(L, w):
    """Observed records of a gradient request: modeling of its true model."""
    obs = w.shots()
    flib.propagate(L, w, Array2D(w.vp_true), Array2D(w.rho_true), obs, 0)
    return obs


def load_observed(w):
    """Observed records of a serving-load gradient request: all-zero traces, so
    the adjoint source is the predicted records themselves.  The work of a
    gradient does not depend on the values of the observed traces, and the
    stream then needs no modeling run per request."""
    return w.shots()


class Done:
    """What is kept of a completed request (its arrays are released)."""

    def __init__(self, job):
        self.problem, self.latency, self.queue = job.problem, job.latency, job.queue
        self.misfit, self.nshot = job.misfit, job.w.nshot
        self.nz, self.nx, self.nt = job.w.nz, job.w.nx, job.w.nt

    def row(self):
        return {"problem": self.problem, "nz": self.nz, "nx": self.nx, "nt": self.nt,
                "shots": self.nshot, "latency_s": self.latency, "queue_s": self.queue}

    def shots(self):
        return self.nshot


def run_stream(server, stream, log=None, ahead=8):
    """Submit every request at its arrival time from a submitter thread (which
    builds each request's workload and arrays shortly before its arrival, at
    most `ahead` requests early) and wait for all of them, releasing each one's
    arrays when it completes.  Returns the completed requests and the wall time
    from the first submission to the last completion."""
    n = len(stream)
    slots = [None] * n
    ready = threading.Semaphore(0)
    built = threading.Semaphore(ahead)
    t0 = time.perf_counter()

    def submitter():
        for k, (ta, problem, p) in enumerate(stream):
            built.acquire()
            w = request_workload(*p)
            obs = load_observed(w) if problem == 1 else None
            job = Job(w, problem, obs)
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            server.submit(job)
            slots[k] = job
            ready.release()

    th = threading.Thread(target=submitter)
    th.start()
    done = []
    for k in range(n):
        ready.acquire()
        job = slots[k]
        server.wait(job)
        done.append(Done(job))
        slots[k] = None
        built.release()
        if log and (k + 1) % max(1, n // 10) == 0:
            log(f"serving: {k + 1}/{n} done")
    th.join()
    return done, time.perf_counter() - t0
