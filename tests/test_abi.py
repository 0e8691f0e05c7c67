"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, its host-side checkpoint bookkeeping matches the oracle's schedule,
and bad arguments end the process the way the reference's err() does -- all
without a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu.records import Array2D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fwiedu.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|long|float|double|char)\s*\**\s*(\w+)\s*\(", txt,
                       flags=re.M)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    L = flib.load()
    names = header_functions()
    assert len(names) >= 29
    assert sorted(names) == sorted(flib.EXPORTED)
    for n in names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.fwiedu_version()


def test_memory_cost_past_32_bit_byte_counts(oracles):
    """calculate_memory_cost_checkpoint at 1000 x 3000 with pml 40 and nt 8000:
    251 checkpoint fields of 3080 x 1080 cells are 3.3e9 bytes, past a 32-bit
    int; the reference forms the byte count as a size_t (checkpoint.c:296)."""
    L = flib.load()
    d, nc, ntp, _ = oracles["f32"].checkpoint_plan(8000)
    want = ntp * 3080 * 1080 * 4 / 1024 ** 3
    assert L.calculate_memory_cost_checkpoint(8000, 3080, 1080) == pytest.approx(want, rel=1e-6)
    assert abs(want - 3.1095) < 1e-3


@pytest.mark.parametrize("nz,nx,pml,want", [(221, 601, 40, 2432), (400, 1200, 40, 768),
                                            (1000, 3000, 40, 128), (2000, 6000, 40, 0)])
def test_lanes_batch_bound(nz, nx, pml, want):
    """The shots-in-lanes layout keeps one field (padded cells x shots x 4 bytes,
    the column height rounded up to 64 cells) below 2 GiB, in whole groups of 64
    shots: the stencil's byte offsets are signed 32-bit.  Closed form against
    the library's bound; 1000 x 3000 holds 128 shots (so its 64-shot workload
    fits in one launch), 2000 x 6000 not even 64."""
    L = flib.load()
    Nx, Nz = nx + 2 * pml, nz + 2 * pml
    ld = (Nz + 63) // 64 * 64
    fbytes = Nx * ld * 4
    n = (2 ** 31 - 1) // fbytes // 64 * 64
    assert n == want
    assert n * fbytes < 2 ** 31 and (n + 64) * fbytes >= 2 ** 31 - 1
    assert L.fwiedu_lanes_max_shots(nz, nx, pml) == want


def test_host_checkpoint_bookkeeping(oracles):
    """alloc_acoustic2dCheckPoint & co. (checkpoint.c:29-299) against the oracle's plan."""
    L = flib.load()
    for nt in (7, 240, 1500, 4001):
        ck = flib.Acoustic2dCheckPoint()
        L.acoustic2dCheckPoint_init(ctypes.byref(ck))
        NZ, NX = 9, 5
        L.alloc_acoustic2dCheckPoint(ctypes.byref(ck), NZ, NX, nt)
        d, nc, ntp, _ = oracles["f32"].checkpoint_plan(nt)
        assert (ck.dCheckPoint, ck.nCheckPoint, ck.ntPoint) == (d, nc, ntp)
        assert ck.memCost == pytest.approx(oracles["f32"].memory_cost_checkpoint(nt, NX, NZ), rel=1e-6)
        assert L.calculate_memory_cost_checkpoint(nt, NX, NZ) == pytest.approx(ck.memCost, rel=1e-7)
        # save a pair at it = d and read it back; status drives the recompute search
        Np = NZ * NX
        p0 = Array2D(np.random.default_rng(0).standard_normal((NX, NZ)).astype(np.float32))
        p1 = Array2D(np.random.default_rng(1).standard_normal((NX, NZ)).astype(np.float32))
        if d < nt:
            L.save_acoustic2dCheckPoint(ctypes.byref(ck), p0.pp, p1.pp, Np, d)
            q0 = Array2D(np.zeros((NX, NZ), np.float32))
            q1 = Array2D(np.zeros((NX, NZ), np.float32))
            L.get_acoustic2dCheckPoint(ctypes.byref(ck), q0.pp, q1.pp, Np, d)
            assert np.array_equal(q0.a, p0.a) and np.array_equal(q1.a, p1.a)
            assert L.get_acoustic2dCheckPointStatus(ctypes.byref(ck), d) == -1
            assert L.get_acoustic2dCheckPointStatus(ctypes.byref(ck), d - 1) == -1
            if d + 1 < nt:
                assert L.get_acoustic2dCheckPointStatus(ctypes.byref(ck), d + 1) == d
        L.free_acoustic2dCheckPoint(ctypes.byref(ck))
        assert not ck.p and not ck.status


def _run_child(code):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "fwi-edu_amd"), ROOT])
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                          timeout=120)


BAD = """
import numpy as np
from fwiedu import lib as flib, synth
from fwiedu.records import Array2D
L = flib.load()
w = synth.workload('tiny')
vp, rho = Array2D(w.vp_true), Array2D(w.rho_true)
obs = w.shots()
L.iso_acoustic2d_propagation(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt,
                             w.nt, {pml}, {fs}, 0.0, 2.0, 1, {problem}, 1, 0, obs.recs, None)
print('returned')
"""


@pytest.mark.parametrize("pml,fs,problem,msg", [
    (2, 0, 0, "pmlThick must be at least 3"),
    (12, 3, 0, "freesurface must be 0 or 1"),
    (12, 0, 5, "problem must be 0, 1 or 2"),
    (12, 0, 1, "problem 1 needs a gradient array"),
])
def test_invalid_arguments_exit_like_err(pml, fs, problem, msg):
    """Argument errors are checked before any device call and exit(1) with a
    message, like the reference's SU err()."""
    r = _run_child(BAD.format(pml=pml, fs=fs, problem=problem))
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert msg in r.stderr
    assert "returned" not in r.stdout


BAD_GEOMETRY = """
import ctypes
import numpy as np
from fwiedu import lib as flib
from fwiedu.records import fptr
L = flib.load()
x = (np.arange(201) * np.float32(10.0)).astype(np.float32)
z = (np.arange(101) * np.float32(10.0)).astype(np.float32)
sx, sz = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_float)()
L.generate_shot_positions(fptr(x), 201, fptr(z), 101, {fsx}, 10.0, {sdepth}, 5,
                          ctypes.byref(sx), ctypes.byref(sz))
print('returned')
"""


@pytest.mark.parametrize("fsx,sdepth,msg", [
    (2001.0, 20.0, "x-coordinate of first shot is out of range"),
    (-10.0, 20.0, "x-coordinate of first shot is out of range"),
    (100.0, 1001.0, "z-coordinate of shots is out of range"),
])
def test_shot_positions_out_of_the_model_exit_like_err(fsx, sdepth, msg):
    """generate_shot_positions (observation.c:138-148) ends the program with
    the reference's messages for shots outside the model."""
    r = _run_child(BAD_GEOMETRY.format(fsx=fsx, sdepth=sdepth))
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert msg in r.stderr
    assert "returned" not in r.stdout


ENGINE_MODE2_UNFILLED = """
import ctypes
import numpy as np
from fwiedu import lib as flib, synth
from fwiedu.records import Array2D
L = flib.load()
w = synth.workload('tiny', nshot=1)
obs = w.shots()
r = obs.recs[0]
nxx = r.ix_max - r.ix_min + 1
vp = Array2D(w.vp_start[r.ix_min:r.ix_max + 1])
rho = Array2D(w.rho_start[r.ix_min:r.ix_max + 1])
ck = flib.Acoustic2dCheckPoint()
L.acoustic2dCheckPoint_init(ctypes.byref(ck))
L.alloc_acoustic2dCheckPoint(ctypes.byref(ck), w.nz + 2 * w.pml, nxx + 2 * w.pml, w.nt)
g = Array2D(np.zeros((nxx, w.nz), np.float32))
L.iso_acoustic2d_propagation_engine(ctypes.byref(r), ctypes.byref(ck), vp.pp, rho.pp, nxx, w.nz,
                                    w.dx, w.dz, w.dt, w.nt, w.pml, w.freesurface, 1, 2, 1, 0, g.pp)
print('returned')
"""


def test_engine_mode2_needs_the_checkpoint_pairs():
    """The per-shot engine's reverse pass restarts from the checkpoint pairs
    whose status mode 1 sets (checkpoint.c:133-156, 236-265): a freshly
    allocated checkpoint (status all 0) ends the call with an error, before
    any device call."""
    r = _run_child(ENGINE_MODE2_UNFILLED)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "the checkpoint holds no field at it=0" in r.stderr
    assert "returned" not in r.stdout
