"""GPU: the verbose debug files of iso_acoustic2d_propagation, its engine and
the device FWI iteration against the strict-order FP32 oracle's
(propagation.c:597-601, 717-723, 917-943, 996-1005).

Same file names; the same bytes for the records, adjoint sources, raw shot
gradients and wavefields (the stencils are bit-exact against the oracle); the
call's raw gradient and pseudo-Hessian within 1e-6 relative L2 (the per-shot
preconditioning divides by a mean that the GPU sums in another order).
"""
import ctypes
import glob
import os
import sys

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu import synth
from fwiedu.records import Array2D

pytestmark = pytest.mark.gpu

SUMMED = ("gradient_raw_iter%03d_%d_%d", "precond_iter%03d_%d_%d")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    lib = flib.load()
    assert lib.fwiedu_device_ok() == 1, "no HIP device"
    return lib


def workload(fs, **kw):
    if fs:
        return synth.workload("tiny", nt=100, freesurface=1, shot_depth_cells=1,
                              rcv_depth_cells=1, **kw)
    return synth.workload("tiny", nt=100, **kw)


def observed(o, w):
    obs = w.shots()
    vp, rho = Array2D(w.vp_true), Array2D(w.rho_true)
    o.propagation(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                  w.freesurface, w.freq, float(w.filter), w.precondition, 0, obs.recs, None)
    return obs


def files(d):
    return {os.path.basename(p): open(p, "rb").read() for p in glob.glob(os.path.join(d, "*.bin"))}


def compare(w, gpu, ora, iter):
    assert set(gpu) == set(ora)
    summed = {f % (iter, w.nz, w.nx) for f in SUMMED}
    for name, b in ora.items():
        if any(name.startswith(p) for p in summed):
            a = np.frombuffer(gpu[name], np.float32).astype(np.float64)
            r = np.frombuffer(b, np.float32).astype(np.float64)
            assert np.linalg.norm(a - r) <= 1e-6 * np.linalg.norm(r), name
        else:
            assert gpu[name] == b, name


def run_in(d, fn, monkeypatch):
    os.makedirs(d, exist_ok=True)
    monkeypatch.chdir(d)
    fn()


@pytest.mark.parametrize("fs", [0, 1])
def test_verbose2_files_match_oracle(L, oracles, tmp_path, monkeypatch, fs):
    say(f"case: verbose 2 gradient and modeling files, tiny nt 100, fs {fs}")
    o = oracles["f32s"]
    w = workload(fs)
    obs = observed(o, w)
    monkeypatch.setenv("shotInterval", "2")
    monkeypatch.setenv("snapInterval", "7")
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    for problem in (1, 0):
        pg, po = obs.header_copy(), obs.header_copy()
        gg = Array2D(np.zeros((w.nx, w.nz), np.float32))
        go = Array2D(np.zeros((w.nx, w.nz), np.float32))
        dg, do = tmp_path / f"gpu{problem}", tmp_path / f"ora{problem}"
        run_in(dg, lambda: flib.propagate(L, w, vp, rho, obs, problem, pg, gg, iter=5, verbose=2),
               monkeypatch)
        run_in(do, lambda: o.propagation_verbose(
            vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
            w.freesurface, w.freq, float(w.filter), w.precondition, problem, po.recs,
            go.pp if problem == 1 else None, 5, 2), monkeypatch)
        fg, fo = files(dg), files(do)
        say(f"  problem {problem}: {len(fg)} files")
        assert len(fo) > 0
        compare(w, fg, fo, 5)


def test_verbose1_lanes_and_session_iterate(L, oracles, tmp_path, monkeypatch):
    """Level 1 through the shots-in-lanes stencils (three shots in one 64-lane
    group) equals the oracle's files; the device FWI iteration with iter 4,
    verbose 1 writes the files of its gradient call, equal to those of
    iso_acoustic2d_propagation with iter 4."""
    say("case: verbose 1 through shots in lanes and fwiedu_session_iterate")
    o = oracles["f32s"]
    w = workload(1)
    obs = observed(o, w)
    monkeypatch.setenv("shotInterval", "1")
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    pg, po = obs.header_copy(), obs.header_copy()
    gg = Array2D(np.zeros((w.nx, w.nz), np.float32))
    go = Array2D(np.zeros((w.nx, w.nz), np.float32))
    L.fwiedu_set_layout(1)
    try:
        run_in(tmp_path / "gpu", lambda: flib.propagate(L, w, vp, rho, obs, 1, pg, gg, iter=4,
                                                        verbose=1), monkeypatch)
        assert flib.stat("layout") == 1
        run_in(tmp_path / "sess", lambda: session_iteration(L, w, obs, 4), monkeypatch)
    finally:
        L.fwiedu_set_layout(-1)
    run_in(tmp_path / "ora", lambda: o.propagation_verbose(
        vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
        w.freesurface, w.freq, float(w.filter), w.precondition, 1, po.recs, go.pp, 4, 1),
        monkeypatch)
    fg, fs_, fo = files(tmp_path / "gpu"), files(tmp_path / "sess"), files(tmp_path / "ora")
    assert len(fo) == 3 * w.nshot + 2
    compare(w, fg, fo, 4)
    assert fs_ == fg


def session_iteration(L, w, obs, iter):
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    sess = L.fwiedu_session_create(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt,
                                   w.nt, w.pml, w.freesurface, w.freq, float(w.filter),
                                   w.precondition)
    it = flib.FwieduIteration()
    it.freq, it.filter = w.freq, float(w.filter)
    it.rz, it.rx = 2, 2
    it.waterdepth = w.waterdepth
    it.vp_water, it.rho_water = w.vp_water, w.rho_water
    it.vpmin, it.vpmax, it.rhomin, it.rhomax = w.vpmin, w.vpmax, w.rhomin, w.rhomax
    it.fraction = 0.01
    it.iter, it.verbose = iter, 1
    L.fwiedu_session_iterate(sess, ctypes.byref(it), None, None, None)
    L.fwiedu_session_destroy(sess)


def test_engine_verbose2_snapshots(L, oracles, tmp_path, monkeypatch):
    """iso_acoustic2d_propagation_engine at verbose 2 names its files after shot
    0 and writes the fields the outer call writes for shot 0."""
    say("case: engine modes 1/2 at verbose 2")
    o = oracles["f32s"]
    w = workload(1, nshot=1)
    obs = observed(o, w)
    monkeypatch.setenv("snapInterval", "9")
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    po = obs.header_copy()
    go = Array2D(np.zeros((w.nx, w.nz), np.float32))
    run_in(tmp_path / "ora", lambda: o.propagation_verbose(
        vp.pp, rho.pp, obs.recs, 1, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml, w.freesurface,
        w.freq, float(w.filter), 0, 1, po.recs, go.pp, 0, 2), monkeypatch)

    def engine():
        pre = obs.header_copy()
        r = pre.recs[0]
        nxx = r.ix_max - r.ix_min + 1
        vpw = Array2D(w.vp_start[r.ix_min:r.ix_max + 1])
        rhw = Array2D(w.rho_start[r.ix_min:r.ix_max + 1])
        ck = flib.Acoustic2dCheckPoint()
        L.acoustic2dCheckPoint_init(ctypes.byref(ck))
        L.alloc_acoustic2dCheckPoint(ctypes.byref(ck), w.nz + 2 * w.pml, nxx + 2 * w.pml, w.nt)
        args = (vpw.pp, rhw.pp, nxx, w.nz, w.dx, w.dz, w.dt, w.nt, w.pml, w.freesurface, 0)
        pc = Array2D(np.zeros((nxx, w.nz), np.float32))
        L.iso_acoustic2d_propagation_engine(ctypes.byref(r), ctypes.byref(ck), *args, 1, 0, 2,
                                            pc.pp)
        L.compute_adjoint_source2d(ctypes.byref(r), ctypes.byref(obs.recs[0]), vp.pp, 0.0, 2,
                                   ctypes.byref(r))
        gg = Array2D(np.zeros((nxx, w.nz), np.float32))
        L.iso_acoustic2d_propagation_engine(ctypes.byref(r), ctypes.byref(ck), *args, 2, 0, 2,
                                            gg.pp)
        L.free_acoustic2dCheckPoint(ctypes.byref(ck))

    run_in(tmp_path / "gpu", engine, monkeypatch)
    fg, fo = files(tmp_path / "gpu"), files(tmp_path / "ora")
    fo = {k: v for k, v in fo.items() if k.startswith("wavefield_")}
    assert len(fo) == 3 * len(range(0, w.nt, 9))
    assert fg == fo
