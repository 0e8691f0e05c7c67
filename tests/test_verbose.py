"""The reference's verbose debug files, restated by the oracle (CPU).

iso_acoustic2d_propagation writes, at verbose >= 1, the predicted records, the
adjoint source and the raw gradient of every shotInterval-th shot of a gradient
and, after the shot loop, the raw gradient and the pseudo-Hessian
(propagation.c:917-943, 996-1005); at verbose >= 2 its engine writes the
forward, reconstructed and adjoint wavefields of those shots every snapInterval
steps (:597-601, 717-723).  These tests pin what the oracle writes against
independent evidence: file names built here from the reference's format
strings, records and gradients of separate oracle calls, the receiver samples
of the forward wavefield, and the imaging condition recomputed in numpy from
the reconstructed and adjoint snapshots.
"""
import glob
import os

import numpy as np
import pytest

from fwiedu import synth
from fwiedu.records import Array2D, fptr

MST = 5


def workload(fs):
    if fs:
        return synth.workload("tiny", nt=100, freesurface=1, shot_depth_cells=1,
                              rcv_depth_cells=1)
    return synth.workload("tiny", nt=100)


def oracle_call(o, w, obs, problem, precondition, iter=0, verbose=0):
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    pre = obs.header_copy()
    g = Array2D(np.zeros((w.nx, w.nz), np.float32))
    o.propagation_verbose(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt,
                          w.pml, w.freesurface, w.freq, float(w.filter), precondition, problem,
                          pre.recs, g.pp if problem == 1 else None, iter, verbose)
    return pre, g.a.copy()


def observed(o, w):
    obs = w.shots()
    vp, rho = Array2D(w.vp_true), Array2D(w.rho_true)
    o.propagation(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                  w.freesurface, w.freq, float(w.filter), w.precondition, 0, obs.recs, None)
    return obs


def expected_names(w, obs, problem, precondition, iter, verbose, shot_iv, snap_iv):
    """The file names the reference's sprintf calls produce."""
    names = set()
    shots = [s for s in range(w.nshot) if s % shot_iv == 0]
    for s in shots:
        r = obs.recs[s]
        nxx = r.ix_max - r.ix_min + 1
        if problem == 1 and verbose >= 1:
            for what in ("predicted", "adjsource"):
                names.add("%s_iter%03d_shot%05d_%d_%d_%.2fms_%.2fm.bin"
                          % (what, iter, s + 1, r.ns, r.ntr, r.dt * 1e3, r.dtr))
            names.add("gradient_raw_iter%03d_shot%05d_%d_%d_%.2fm_%.2fm.bin"
                      % (iter, s + 1, w.nz, nxx, w.dz, w.dx))
        if verbose >= 2:
            Nz, Nx = w.nz + 2 * w.pml, nxx + 2 * w.pml
            kinds = ["wavefield_forward"]
            if problem == 1:
                kinds += ["wavefield_forward_reconstruction", "wavefield_backward"]
            for it in range(0, w.nt, snap_iv):
                for k in kinds:
                    names.add("%s_iter%03d_ishot%04d_it%05d_%d_%d_%.2fm_%.2fm.bin"
                              % (k, iter, s, it, Nz, Nx, w.dz, w.dx))
    if problem == 1 and verbose >= 1:
        names.add("gradient_raw_iter%03d_%d_%d_%.2fm_%.2fm.bin" % (iter, w.nz, w.nx, w.dz, w.dx))
        if precondition == 1:
            names.add("precond_iter%03d_%d_%d_%.2fm_%.2fm.bin" % (iter, w.nz, w.nx, w.dz, w.dx))
    return names


def load(path, shape=None):
    a = np.fromfile(path, dtype=np.float32)
    return a.reshape(shape) if shape else a


@pytest.mark.parametrize("fs", [0, 1])
def test_oracle_verbose_files(oracles, tmp_path, monkeypatch, fs):
    o = oracles["f32s"]
    w = workload(fs)
    obs = observed(o, w)
    monkeypatch.setenv("shotInterval", "2")
    monkeypatch.setenv("snapInterval", "1")
    monkeypatch.chdir(tmp_path)
    pre, grad = oracle_call(o, w, obs, 1, 1, iter=7, verbose=2)
    names = {os.path.basename(p) for p in glob.glob(str(tmp_path / "*.bin"))}
    assert names == expected_names(w, obs, 1, 1, 7, 2, 2, 1)

    # predicted = the modeling records of the same model; adjsource = what the
    # call leaves in pre
    model, _ = oracle_call(o, w, obs, 0, 1)
    for s in (0, 2):
        r = obs.recs[s]
        tag = "_iter007_shot%05d_%d_%d_%.2fms_%.2fm.bin" % (s + 1, r.ns, r.ntr, r.dt * 1e3, r.dtr)
        assert np.array_equal(load("predicted" + tag, (r.ntr, r.ns)), model.trace(s))
        assert np.array_equal(load("adjsource" + tag, (r.ntr, r.ns)), pre.trace(s))

    # the raw gradient is the one before the global preconditioning
    tag = "_iter007_%d_%d_%.2fm_%.2fm.bin" % (w.nz, w.nx, w.dz, w.dx)
    raw = load("gradient_raw" + tag)
    pc = load("precond" + tag)
    g2 = raw.copy()
    o.global_precondition(g2.ctypes.data_as(o.global_precondition.argtypes[0]),
                          pc.ctypes.data_as(o.global_precondition.argtypes[1]), g2.size, 0.005)
    assert np.array_equal(g2, grad.ravel())

    for s in (0, 2):
        r = obs.recs[s]
        nxx = r.ix_max - r.ix_min + 1
        Nz, Nx = w.nz + 2 * w.pml, nxx + 2 * w.pml

        def snap(kind, it):
            return load("%s_iter007_ishot%04d_it%05d_%d_%d_%.2fm_%.2fm.bin"
                        % (kind, s, it, Nz, Nx, w.dz, w.dx), (Nx, Nz))

        # forward wavefield of level it: the receivers sample it at step it
        igx = np.ctypeslib.as_array(r.igx, (r.ntr,)) + w.pml
        igz = np.ctypeslib.as_array(r.igz, (r.ntr,)) + w.pml
        rec = model.trace(s)
        for it in range(w.nt):
            u = snap("wavefield_forward", it)
            assert np.array_equal(u[igx, igz], rec[:, it]), it
            if fs and it > 0:
                # the free surface's mirror (propagation.c:94-100), all columns
                for i in range(w.pml + 1):
                    assert np.array_equal(u[:, w.pml - i], -u[:, w.pml + i])
        assert not snap("wavefield_forward", 0).any()

        # the reconstructed field at imaging steps is u(it) recomputed, within
        # FP32 rounding (the source in float: T3) without a free surface; with
        # one, the reconstruction adds the source after the mirror (T2), so a
        # source one cell deep leaves the mirror row without it and the two
        # fields part (the imaging check below covers that case exactly)
        for it in range(MST, w.nt, MST):
            a, b = snap("wavefield_forward_reconstruction", it), snap("wavefield_forward", it)
            if not fs:
                assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), it
            assert np.linalg.norm(a - b) <= 0.5 * np.linalg.norm(b), it

        # the imaging condition (propagation.c:706-712, 740) recomputed from
        # the adjoint snapshots (p1 of step it, p0 = p1 of step it + 1, p2 =
        # p1 of step it - 1 before its injection, equal away from the
        # receivers) and the reconstructed ones equals the raw shot gradient
        # at every interior cell without a receiver
        dt = np.float32(w.dt)
        invdtdt = np.float32(1.0) / (dt * dt)
        acc = np.zeros((nxx, w.nz))
        sl = (slice(w.pml, w.pml + nxx), slice(w.pml, w.pml + w.nz))
        for it in range(w.nt - 1, 0, -1):
            if it % MST:
                continue
            p2 = snap("wavefield_backward", it - 1)[sl]
            p1 = snap("wavefield_backward", it)[sl]
            p0 = snap("wavefield_backward", it + 1)[sl] if it + 1 < w.nt else 0 * p1
            u = snap("wavefield_forward_reconstruction", it)[sl]
            acc += ((p2 - np.float32(2.0) * p1 + p0) * invdtdt * u).astype(np.float64)
        vw = w.vp_start[r.ix_min:r.ix_max + 1]
        rw = w.rho_start[r.ix_min:r.ix_max + 1]
        g = (2.0 * acc / (rw * vw).astype(np.float64)).astype(np.float32)
        gf = load("gradient_raw_iter007_shot%05d_%d_%d_%.2fm_%.2fm.bin"
                  % (s + 1, w.nz, nxx, w.dz, w.dx), (nxx, w.nz))
        mask = np.ones((nxx, w.nz), bool)
        mask[igx - w.pml, igz - w.pml] = False
        assert mask.sum() > 0.9 * mask.size
        assert np.array_equal(g[mask], gf[mask])

        # one step of the adjoint recursion (propagation.c:676-699): the adjoint
        # p1 of step it - 1 is the stencil of steps it, it + 1, the free
        # surface, then the injection of the adjoint source of step it - 1 at
        # the receivers (so the snapshot of a step holds its injection)
        co = {k: np.zeros((Nx, Nz), np.float32) for k in
              ("D", "Beta", "c1", "c2", "c3", "K", "ir", "irx", "irz")}
        o.setup_coefs(fptr(np.ascontiguousarray(vw)), fptr(np.ascontiguousarray(rw)), nxx, w.nz,
                      w.pml, w.dx, w.dz, w.dt, r.fpeak,
                      *[fptr(co[k]) for k in ("D", "Beta", "c1", "c2", "c3", "K", "ir", "irx",
                                              "irz")])
        dx32, dz32 = np.float32(w.dx), np.float32(w.dz)
        invdx, invdz = np.float32(1) / dx32, np.float32(1) / dz32
        izs = w.pml if fs else 3
        adj = pre.trace(s)
        for it in (w.nt - 2, 50, 11, 1):
            p1 = np.ascontiguousarray(snap("wavefield_backward", it))
            p0 = np.ascontiguousarray(snap("wavefield_backward", it + 1))
            p2 = np.zeros((Nx, Nz), np.float32)
            o.fd_step(fptr(p2), fptr(p1), fptr(p0), fptr(co["ir"]), fptr(co["irx"]),
                      fptr(co["irz"]), fptr(co["K"]), fptr(co["c1"]), fptr(co["c2"]),
                      fptr(co["c3"]), Nx, Nz, invdx, invdz, np.float32(1) / (dx32 * dx32),
                      np.float32(1) / (dz32 * dz32), 3, Nx - 4, izs, Nz - 4)
            if fs:
                o.freesurface(fptr(p2), Nx, Nz, w.pml)
            for k in range(r.ntr):
                term = (adj[k, it - 1] * co["K"][igx[k], igz[k]] * dt * dt * invdx * invdz)
                p2[igx[k], igz[k]] -= np.float32(term)
            assert np.array_equal(p2, snap("wavefield_backward", it - 1)), it


def test_oracle_verbose_levels(oracles, tmp_path, monkeypatch):
    """verbose 1 writes no wavefields; modeling writes only wavefields (level
    2); without precondition no pseudo-Hessian; the raw gradient then equals
    the output and the shot-order sum of the raw shot gradients."""
    o = oracles["f32s"]
    w = workload(0)
    obs = observed(o, w)
    monkeypatch.setenv("shotInterval", "1")
    monkeypatch.setenv("snapInterval", "30")
    d1 = tmp_path / "g"
    d1.mkdir()
    monkeypatch.chdir(d1)
    _, grad = oracle_call(o, w, obs, 1, 0, iter=2, verbose=1)
    names = {os.path.basename(p) for p in glob.glob("*.bin")}
    assert names == expected_names(w, obs, 1, 0, 2, 1, 1, 30)
    raw = load("gradient_raw_iter002_%d_%d_%.2fm_%.2fm.bin" % (w.nz, w.nx, w.dz, w.dx))
    assert np.array_equal(raw, grad.ravel())
    tot = np.zeros((w.nx, w.nz))
    for s in range(w.nshot):
        r = obs.recs[s]
        nxx = r.ix_max - r.ix_min + 1
        gs = load("gradient_raw_iter002_shot%05d_%d_%d_%.2fm_%.2fm.bin"
                  % (s + 1, w.nz, nxx, w.dz, w.dx), (nxx, w.nz))
        tot[r.ix_min:r.ix_max + 1] += gs
    assert np.array_equal(tot.astype(np.float32), grad)

    d2 = tmp_path / "m"
    d2.mkdir()
    monkeypatch.chdir(d2)
    oracle_call(o, w, obs, 0, 1, iter=0, verbose=2)
    names = {os.path.basename(p) for p in glob.glob("*.bin")}
    assert names == expected_names(w, obs, 0, 1, 0, 2, 1, 30)
