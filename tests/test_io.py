"""Data formats of the drivers (host code of the library, no GPU): JSON
acquisition geometry, grid geometry, peak frequency, SU shot gathers.

Pins:
  * JSON: the reference demo's acquisition file (first 3 of its 30 shots,
    tests/golden/acquisition_3shots.json) read by the library against Python's
    json module; ns, dt and the mean receiver spacing of init_observe2d_from_JSON
    (observation.c:509-517); fpeak of the reference's shipped Ricker wavelet
    (bin 60 of 6001 samples at 1 ms: 9.99833 Hz).
  * geometry from parameters (generate_shot_positions, init_observe2d,
    observation.c:102-165, 229-336): the demo's acquisition file is the
    fixed full spread of its modeling.sh parameters (offsmin 0, offsmax 600,
    dtr 1, gdepth 2 cells at 12.5 m) with shots every 20 cells from cell 0 at
    depth 2: the parameter path reproduces its first 3 shots exactly; shot
    counts cut at the model's edge; a spread cut by the model's edge; the
    reference loops' dropped slot when offsmax is not a multiple of dtr.
  * grid geometry: floorf(coord / d) in float and the +-30 window
    (observation.c:376-418), recomputed with numpy float32.
  * SU: the byte positions of SEG-Y rev 1 (bytes 0-179: tracl 0, tracr 4,
    fldr 8, tracf 12, cdp 20, trid 28, offset 36, gelev 40, sdepth 48,
    scalel 68, scalco 70, sx 72, gx 80, ns 114, dt 116); the Seismic Unix
    extension (d1 180, d2 188, ntr 204) follows SU's segy.h, which is not in
    this image: parity unpinned for those three fields beyond this layout.
    The file name format is pinned by the demo's gather name
    shots_p_30_3001_18030_02.00ms.su (fwi.sh:43).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu import synth
from fwiedu.records import Array2D, Record2D, fptr

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def L():
    return flib.load()


def ricker():
    return np.fromfile(os.path.join(GOLDEN, "ricker_6001_1ms_10Hz_delay0.15ms.bin"), np.float32)


def test_peak_frequency_of_the_shipped_wavelet(L):
    src = ricker()
    f = L.calculate_peak_frequency(fptr(src), src.size, 0.001)
    assert f == np.float32(60) * (np.float32(1.0) / (np.float32(6001) * np.float32(0.001)))
    assert abs(f - 9.99833) < 1e-4


def json_records(L, sample_rate=2):
    with open(os.path.join(GOLDEN, "acquisition_3shots.json")) as fh:
        d = json.load(fh)
    m = d["model"]
    nx, nz, dx, dz = m["nx"], m["nz"], m["dx"], m["dz"]
    x = (np.arange(nx) * dx).astype(np.float32)
    z = (np.arange(nz) * dz).astype(np.float32)
    src = ricker()
    recs = (Record2D * 3)()
    L.fwiedu_observe2d_from_json(recs, os.path.join(GOLDEN, "acquisition_3shots.json").encode(), 3,
                                 fptr(x), nx, fptr(z), nz, 0.001, src.size, fptr(src), sample_rate)
    return d, recs, (nx, nz, dx, dz)


def test_json_geometry(L):
    d, recs, (nx, nz, dx, dz) = json_records(L)
    for s, js in enumerate(d["record"]):
        r = recs[s]
        assert r.shotID == js["shotID"] and r.ntr == js["ntr"] == 601
        assert r.sx == np.float32(js["sx"]) and r.sz == np.float32(js["sz"])
        gx = np.ctypeslib.as_array(r.gx, (r.ntr,))
        gz = np.ctypeslib.as_array(r.gz, (r.ntr,))
        assert np.array_equal(gx, np.asarray(js["gx"], np.float32))
        assert np.array_equal(gz, np.asarray(js["gz"], np.float32))
        assert r.sampleRate == 2 and r.ns == 3001 and r.dt == np.float32(0.001) * 2
        assert abs(r.dtr - 12.5) < 1e-5
        assert abs(r.fpeak - 9.99833) < 1e-4
    assert [recs[s].sx for s in range(3)] == [0.0, 250.0, 500.0]


def param_records(L, nx, nz, dx, dz, fsx, ds, sdepth, nshot, offsmin, offsmax, dtr, gdepth,
                   sample_rate=1):
    """The modeling driver's parameter geometry (cells times dx / dz, as
    acoustic2d_modeling.c:240-246) through the library."""
    f32 = np.float32
    x = (np.arange(nx) * f32(dx)).astype(f32)
    z = (np.arange(nz) * f32(dz)).astype(f32)
    src = ricker()
    sx, sz = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_float)()
    n = L.generate_shot_positions(fptr(x), nx, fptr(z), nz, f32(fsx) * f32(dx), f32(ds) * f32(dx),
                                  f32(sdepth) * f32(dz), nshot, ctypes.byref(sx), ctypes.byref(sz))
    recs = (Record2D * n)()
    L.init_observe2d(recs, n, fptr(x), nx, fptr(z), nz, 0.001, src.size, sx, sz, fptr(src),
                     f32(offsmin) * f32(dx), f32(offsmax) * f32(dx), f32(dtr) * f32(dx),
                     f32(gdepth) * f32(dz), sample_rate)
    return n, recs


def test_parameter_geometry_reproduces_the_demo_acquisition(L):
    d, jrecs, (nx, nz, dx, dz) = json_records(L, 2)
    n, recs = param_records(L, nx, nz, dx, dz, fsx=0, ds=20, sdepth=2, nshot=3, offsmin=0,
                            offsmax=600, dtr=1, gdepth=2, sample_rate=2)
    assert n == 3
    for s in range(3):
        a, b = recs[s], jrecs[s]
        for k in ("shotID", "sx", "sz", "ntr", "sampleRate", "ns", "dt", "dtr", "fpeak"):
            assert getattr(a, k) == getattr(b, k), (s, k)
        for k in ("gx", "gz"):
            assert np.array_equal(np.ctypeslib.as_array(getattr(a, k), (a.ntr,)),
                                  np.ctypeslib.as_array(getattr(b, k), (b.ntr,))), (s, k)


def test_parameter_geometry_edges(L):
    f32 = np.float32
    # the reference's defaults (fsx 1, ds 1, 99999 shots) on 201 columns: the
    # shots stop at the last column, 200 of them
    n, recs = param_records(L, 201, 101, 10.0, 10.0, fsx=1, ds=1, sdepth=2, nshot=99999,
                            offsmin=0, offsmax=99999, dtr=1, gdepth=2)
    assert n == 200
    assert [recs[i].sx for i in (0, 1, 199)] == [10.0, 20.0, 2000.0]
    assert all(recs[i].ntr == 201 and recs[i].shotID == i + 1 for i in range(n))
    # a spread of offsets 0..50 cells each side, cut by the model's edges
    n, recs = param_records(L, 201, 101, 10.0, 10.0, fsx=30, ds=70, sdepth=2, nshot=5,
                            offsmin=0, offsmax=50, dtr=1, gdepth=3)
    assert n == 3 and [recs[i].sx for i in range(3)] == [300.0, 1000.0, 1700.0]
    for i, lo, hi in ((0, 0, 80), (1, 50, 150), (2, 120, 200)):
        gx = np.ctypeslib.as_array(recs[i].gx, (recs[i].ntr,))
        assert np.array_equal(gx, (np.arange(lo, hi + 1) * f32(10.0)).astype(f32)), i
        assert np.all(np.ctypeslib.as_array(recs[i].gz, (recs[i].ntr,)) == f32(30.0))
        assert recs[i].dtr == f32(10.0)
    # offsets 0..2.5 cells in steps of 1: the first loop gives -2.5, -1.5,
    # -0.5 and its last slot goes to the zero offset (observation.c:259-272),
    # so the spread is -2.5, -1.5, 0, 1, 2 cells from the shot
    n, recs = param_records(L, 201, 101, 10.0, 10.0, fsx=100, ds=1, sdepth=2, nshot=1,
                            offsmin=0, offsmax=2.5, dtr=1, gdepth=2)
    gx = np.ctypeslib.as_array(recs[0].gx, (recs[0].ntr,))
    assert np.array_equal(gx, np.array([975, 985, 1000, 1010, 1020], f32))


def test_create_grid_observe2d(L):
    d, recs, (nx, nz, dx, dz) = json_records(L, 1)
    L.create_grid_observe2d(recs, 3, dx, dz, nx)
    f32 = np.float32
    for s, js in enumerate(d["record"]):
        r = recs[s]
        igx = np.floor(np.asarray(js["gx"], f32) / f32(dx)).astype(np.int64)
        igz = np.floor(np.asarray(js["gz"], f32) / f32(dz)).astype(np.int64)
        ix_min, ix_max = max(igx.min() - 30, 0), min(igx.max() + 30, nx - 1)
        assert (r.ix_min, r.ix_max) == (ix_min, ix_max)
        assert r.isx == int(np.floor(f32(js["sx"]) / f32(dx))) - ix_min
        assert r.isz == int(np.floor(f32(js["sz"]) / f32(dz)))
        assert np.array_equal(np.ctypeslib.as_array(r.igx, (r.ntr,)), igx - ix_min)
        assert np.array_equal(np.ctypeslib.as_array(r.igz, (r.ntr,)), igz)


def test_su_round_trip_and_header_layout(L, tmp_path):
    w = synth.workload("tiny", nshot=2, sample_rate=2)
    shots = w.shots()
    rng = np.random.default_rng(3)
    for s in range(w.nshot):
        shots.set_traces(s, rng.standard_normal(shots.trace(s).shape).astype(np.float32))
    prefix = str(tmp_path / "shots_p")
    L.write_acoustic2d_traces(prefix.encode(), shots.recs, w.nshot)
    r0 = shots.recs[0]
    ntr_tot = sum(shots.recs[s].ntr for s in range(w.nshot))
    name = f"{prefix}_{w.nshot}_{r0.ns}_{ntr_tot}_{r0.dt * 1000:05.2f}ms.su"
    assert os.path.exists(name), os.listdir(tmp_path)
    raw = np.fromfile(name, np.uint8)
    ns = r0.ns
    assert raw.size == ntr_tot * (240 + 4 * ns)

    def field(k, off, dtype):
        base = k * (240 + 4 * ns) + off
        return np.frombuffer(raw[base:base + np.dtype(dtype).itemsize].tobytes(), dtype)[0]

    k = 0
    for s in range(w.nshot):
        r = shots.recs[s]
        for i in range(r.ntr):
            assert field(k, 0, "<i4") == k + 1 and field(k, 4, "<i4") == k + 1   # tracl, tracr
            assert field(k, 8, "<i4") == r.shotID and field(k, 12, "<i4") == i + 1  # fldr, tracf
            assert field(k, 28, "<i2") == 1                                      # trid
            assert field(k, 68, "<i2") == -1000 and field(k, 70, "<i2") == -1000  # scalel, scalco
            sx = int(np.float32(r.sx) * np.float32(1000.0))
            gx = int(np.float32(r.gx[i]) * np.float32(1000.0))
            assert field(k, 72, "<i4") == sx and field(k, 80, "<i4") == gx
            assert field(k, 36, "<i4") == gx - sx                                # offset
            assert field(k, 20, "<i4") == int((gx + sx) / 2)                     # cdp
            assert field(k, 48, "<i4") == int(np.float32(r.sz) * np.float32(1000.0))
            assert field(k, 40, "<i4") == int(np.float32(r.gz[i]) * np.float32(1000.0))
            assert field(k, 114, "<u2") == ns
            assert field(k, 116, "<u2") == int(np.float32(r.dt) * np.float32(1e6))
            assert field(k, 180, "<f4") == np.float32(r.dt) and field(k, 188, "<f4") == r.dtr
            assert field(k, 204, "<i4") == r.ntr
            base = k * (240 + 4 * ns) + 240
            assert np.array_equal(np.frombuffer(raw[base:base + 4 * ns].tobytes(), "<f4"),
                                  shots.trace(s)[i])
            k += 1
    assert L.get_shots_number(name.encode()) == w.nshot
    back = (Record2D * w.nshot)()
    L.read_gather2d_from_sufile(back, name.encode(), w.nshot, w.nt)
    for s in range(w.nshot):
        a, b = shots.recs[s], back[s]
        assert b.shotID == s + 1 and b.ntr == a.ntr and b.ns == a.ns
        assert b.sampleRate == (w.nt - 1) // (a.ns - 1) == 2
        assert b.dt == np.float32(np.uint16(int(np.float32(a.dt) * np.float32(1e6)))) / np.float32(1e6)
        assert abs(b.sx - a.sx) <= 1e-3 and abs(b.sz - a.sz) <= 1e-3
        tr = np.ctypeslib.as_array(b.traces_p[0], (b.ntr * b.ns,)).reshape(b.ntr, b.ns)
        assert np.array_equal(tr, shots.trace(s))
        gx = np.ctypeslib.as_array(b.gx, (b.ntr,))
        assert np.max(np.abs(gx - np.ctypeslib.as_array(a.gx, (a.ntr,)))) <= 1e-3
        assert abs(b.dtr - a.dtr) <= 1e-6 * abs(a.dtr)


def test_su_file_name_of_the_demo_gather(L, tmp_path):
    """30 shots x 601 traces x 3001 samples at 2 ms -> the demo's
    shots_p_30_3001_18030_02.00ms.su (one shared trace buffer, so only the file
    is large)."""
    ns, ntr = 3001, 601
    buf = np.zeros(ns, np.float32)
    rows = (ctypes.POINTER(ctypes.c_float) * ntr)(*([fptr(buf)] * ntr))
    gx = np.arange(ntr, dtype=np.float32) * np.float32(12.5)
    gz = np.full(ntr, 25.0, np.float32)
    recs = (Record2D * 30)()
    for s in range(30):
        r = recs[s]
        r.shotID, r.sx, r.sz, r.ntr, r.ns = s + 1, 250.0 * s, 25.0, ntr, ns
        r.dt, r.dtr = 0.002, 12.5
        r.gx, r.gz = fptr(gx), fptr(gz)
        r.traces_p = ctypes.cast(rows, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)))
    prefix = str(tmp_path / "shots_p")
    L.write_acoustic2d_traces(prefix.encode(), recs, 30)
    assert os.listdir(tmp_path) == ["shots_p_30_3001_18030_02.00ms.su"]
    assert L.get_shots_number(str(tmp_path / "shots_p_30_3001_18030_02.00ms.su").encode()) == 30
