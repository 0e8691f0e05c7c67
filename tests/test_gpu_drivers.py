"""The C drivers (fwi-edu_amd/bin/acoustic2d_modeling, acoustic2d_fwi), built
on the library's C ABI, run as separate processes on files:

  * modeling: float32 model files, wavelet file and a JSON acquisition in,
    SU gather out; its traces are bit-identical to iso_acoustic2d_propagation
    called in-process on the same model and geometry, and the same acquisition
    given by the geometry parameters (fsx, ds, sdepth, offsmin, offsmax, dtr,
    gdepth) writes the same file;
  * FWI: that gather, a smoothed start model and one iteration; the written
    model and gradient are bit-identical to fwiedu_session_iterate called
    in-process on the same inputs, and the final predicted gather is written;
  * verbose=1: both drivers write their input models and wavelet
    (acoustic2d_modeling.c:327-336, acoustic2d_fwi.c:370-381), and the FWI
    driver's gradient writes the library's per-shot debug files, the same
    bytes as the in-process iteration with iter 1, verbose 1.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu import synth
from fwiedu.records import Array2D, Record2D

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fwi-edu_amd", "bin")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    lib = flib.load()
    assert lib.fwiedu_device_ok() == 1, "no HIP device"
    return lib


def run(args, cwd):
    say("  $", " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)))
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stdout


def records_from_json(L, path, w, x, z):
    n = L.fwiedu_json_shot_count(path.encode())
    recs = (Record2D * n)()
    L.fwiedu_observe2d_from_json(recs, path.encode(), n, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 w.nx, z.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.nz, w.dt,
                                 w.nt, w.src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1)
    L.create_grid_observe2d(recs, n, w.dx, w.dz, w.nx)
    return recs, n


def traces(rec):
    return np.ctypeslib.as_array(rec.traces_p[0], (rec.ntr * rec.ns,)).copy()


def test_modeling_and_fwi_drivers(L, tmp_path, monkeypatch):
    w = synth.workload("tiny", nshot=3, freesurface=1, shot_depth_cells=1)
    say("case: C drivers, tiny 30x48, 3 shots")
    d = str(tmp_path)
    w.vp_true.astype(np.float32).tofile(os.path.join(d, "vp.bin"))
    w.rho_true.astype(np.float32).tofile(os.path.join(d, "rho.bin"))
    w.vp_start.astype(np.float32).tofile(os.path.join(d, "vp0.bin"))
    w.rho_start.astype(np.float32).tofile(os.path.join(d, "rho0.bin"))
    w.src.astype(np.float32).tofile(os.path.join(d, "src.bin"))
    gx = [float(np.float32(i) * np.float32(w.dx)) for i in range(w.nx)]
    acq = {"model": {"nx": w.nx, "nz": w.nz, "dx": w.dx, "dz": w.dz},
           "record": [{"shotID": s + 1, "sx": float(c * w.dx), "sz": float(w.dz), "ntr": w.nx,
                       "gx": gx, "gz": [float(w.dz)] * w.nx}
                      for s, c in enumerate(w.shot_columns())]}
    with open(os.path.join(d, "acq.json"), "w") as f:
        json.dump(acq, f)
    common = [f"nx={w.nx}", f"nz={w.nz}", f"nt={w.nt}", f"dx={w.dx}", f"dz={w.dz}", f"dt={w.dt}",
              f"pmlThick={w.pml}", "freesurface=1", "Gardner=0", "srcfile=src.bin"]
    run([os.path.join(BIN, "acoustic2d_modeling")] + common +
        ["vpfile=vp.bin", "rhofile=rho.bin", "jsonfile=acq.json", "verbose=1"], d)
    su = [f for f in os.listdir(d) if f.startswith("shots_p_") and f.endswith(".su")]
    assert len(su) == 1, os.listdir(d)
    mtag = f"{w.nz}_{w.nx}_{w.dz:.2f}m_{w.dx:.2f}m.bin"
    wav = f"wavelet_{w.nt}_{w.dt * 1000:.2f}ms.bin"
    assert np.array_equal(np.fromfile(os.path.join(d, "vp_" + mtag), np.float32),
                          w.vp_true.ravel())
    assert np.array_equal(np.fromfile(os.path.join(d, "rho_" + mtag), np.float32),
                          w.rho_true.ravel())
    assert np.array_equal(np.fromfile(os.path.join(d, wav), np.float32), w.src)

    # in-process modeling on the same model and geometry
    x = (np.arange(w.nx) * np.float32(w.dx)).astype(np.float32)
    z = (np.arange(w.nz) * np.float32(w.dz)).astype(np.float32)
    recs, n = records_from_json(L, os.path.join(d, "acq.json"), w, x, z)
    vp, rho = Array2D(w.vp_true), Array2D(w.rho_true)
    L.iso_acoustic2d_propagation(vp.pp, rho.pp, recs, n, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                                 1, 0.0, 0.0, 0, 0, 0, 0, recs, None)
    back = (Record2D * n)()
    L.read_gather2d_from_sufile(back, os.path.join(d, su[0]).encode(), n, w.nt)
    for s in range(n):
        assert np.array_equal(traces(back[s]), traces(recs[s])), s

    # the same acquisition from the reference's geometry parameters (shots
    # every 16 cells from cell 8 at depth 1, the full spread at depth 1): the
    # same gather file, byte for byte
    cols = list(w.shot_columns())
    assert cols == [8, 24, 40]
    dp = os.path.join(d, "params")
    os.mkdir(dp)
    run([os.path.join(BIN, "acoustic2d_modeling")] + common +
        ["vpfile=../vp.bin", "rhofile=../rho.bin", "srcfile=../src.bin", "Nshot=3", "fsx=8",
         "ds=16", "sdepth=1", "offsmin=0", f"offsmax={w.nx - 1}", "dtr=1", "gdepth=1"], dp)
    with open(os.path.join(d, su[0]), "rb") as fa, open(os.path.join(dp, su[0]), "rb") as fb:
        assert fa.read() == fb.read()

    # FWI driver, one iteration, against fwiedu_session_iterate in-process
    run([os.path.join(BIN, "acoustic2d_fwi")] + common +
        ["vpfile=vp0.bin", "rhofile=rho0.bin", f"recordfile={su[0]}", f"Nshot={n}",
         "iteration=1", "freq=0", "filter=2", "rx=3", "rz=3", "precondition=1",
         f"waterdepth={w.waterdepth}", "vpmin=1400", "vpmax=5000", "rhomin=1000", "rhomax=3000",
         "verbose=1"], d)
    tag = f"{w.nz:05d}_{w.nx:05d}_{w.dz:.2f}m_{w.dx:.2f}m"
    vp_drv = np.fromfile(os.path.join(d, f"vp_iter001_{tag}.bin"), np.float32).reshape(w.nx, w.nz)
    g_drv = np.fromfile(os.path.join(d, f"grad_iter001_{tag}.bin"), np.float32).reshape(w.nx, w.nz)
    assert any(f.startswith("shots_p_predicted_") for f in os.listdir(d))

    obs = (Record2D * n)()
    L.read_gather2d_from_sufile(obs, os.path.join(d, su[0]).encode(), n, w.nt)
    fpeak = L.calculate_peak_frequency(w.src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.nt, w.dt)
    for s in range(n):
        obs[s].src = w.src.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        obs[s].fpeak = fpeak
    L.create_grid_observe2d(obs, n, w.dx, w.dz, w.nx)
    vp0, rho0 = Array2D(w.vp_start), Array2D(w.rho_start)
    vp0.a[:, :w.waterdepth] = 1500.0
    rho0.a[:, :w.waterdepth] = 1000.0
    sess = L.fwiedu_session_create(vp0.pp, rho0.pp, obs, n, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt,
                                   w.pml, 1, 0.0, 2.0, 1)
    p = flib.FwieduIteration(freq=0.0, filter=2.0, rz=3, rx=3, waterdepth=w.waterdepth,
                             vp_water=1500.0, rho_water=1000.0, vpmin=1400.0, vpmax=5000.0,
                             rhomin=1000.0, rhomax=3000.0, gardner=0, fraction=0.01, iter=1,
                             verbose=1)
    vg, rg, gg = (Array2D(np.zeros((w.nx, w.nz), np.float32)) for _ in range(3))
    dv = tmp_path / "inproc"
    dv.mkdir()
    monkeypatch.chdir(dv)
    try:
        L.fwiedu_session_iterate(sess, ctypes.byref(p), vg.pp, rg.pp, gg.pp)
    finally:
        L.fwiedu_session_destroy(sess)
    say(f"  fun1 {p.fun1:.6e} step {p.steplen:.6e}")
    assert np.array_equal(vp_drv, vg.a)
    assert np.array_equal(g_drv, gg.a)

    # verbose=1: the start model as the driver holds it after the water rows,
    # and the gradient's debug files (shot 1 of 3 at shotInterval 20)
    assert np.array_equal(np.fromfile(os.path.join(d, "vp_init_" + mtag), np.float32),
                          vp0.a.ravel())
    assert np.array_equal(np.fromfile(os.path.join(d, "rho_init_" + mtag), np.float32),
                          rho0.a.ravel())
    assert np.array_equal(np.fromfile(os.path.join(d, wav), np.float32), w.src)
    dbg = sorted(f for f in os.listdir(dv) if f.endswith(".bin"))
    r = obs[0]
    nxx = r.ix_max - r.ix_min + 1
    assert dbg == sorted([
        "adjsource_iter001_shot00001_%d_%d_%.2fms_%.2fm.bin" % (r.ns, r.ntr, r.dt * 1e3, r.dtr),
        "predicted_iter001_shot00001_%d_%d_%.2fms_%.2fm.bin" % (r.ns, r.ntr, r.dt * 1e3, r.dtr),
        "gradient_raw_iter001_shot00001_%d_%d_%.2fm_%.2fm.bin" % (w.nz, nxx, w.dz, w.dx),
        "gradient_raw_iter001_" + mtag, "precond_iter001_" + mtag])
    for f in dbg:
        with open(os.path.join(d, f), "rb") as fa, open(os.path.join(dv, f), "rb") as fb:
            assert fa.read() == fb.read(), f
