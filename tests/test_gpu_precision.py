"""Checkpoint storage precision: FP32 against FP16 with a per-snapshot
power-of-two scale, each stated against the FP64 oracle.

Arithmetic is FP32 in both; FP16 storage only changes the fields the reverse
pass restarts from (and images at checkpoint times), so
  * records and adjoint sources are bit-identical between the two storages;
  * the FP16 gradient stays within 1e-2 relative L2 of the FP64 oracle's
    (FP16 keeps 11 significant bits, 4.9e-4 relative per stored value);
  * the FP32 gradient stays within 1e-3 of the FP64 oracle's, as in the parity
    tests.
The measured figures are printed (pytest -s) and quoted in the design notes.
"""
import sys

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu import synth
from fwiedu.records import Array2D

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    lib = flib.load()
    assert lib.fwiedu_device_ok() == 1, "no HIP device"
    return lib


def observed(o, w):
    obs = w.shots()
    vp, rho = Array2D(w.vp_true), Array2D(w.rho_true)
    o.propagation(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                  w.freesurface, w.freq, float(w.filter), w.precondition, 0, obs.recs, None)
    return obs


def gpu_gradient(L, w, obs, bits):
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    pre = obs.header_copy()
    g = Array2D(np.zeros((w.nx, w.nz), np.float32))
    L.fwiedu_set_checkpoint_bits(bits)
    try:
        flib.propagate(L, w, vp, rho, obs, 1, pre, g)
        assert flib.stat("ckpt_bits") == bits
    finally:
        L.fwiedu_set_checkpoint_bits(32)
    return pre, g.a.copy()


def recs(shotset):
    return np.concatenate([t.ravel() for t in shotset.all_traces()])


CASES = {
    "tiny_fs1": dict(name="tiny", freesurface=1, shot_depth_cells=1, rcv_depth_cells=1),
    "c0": dict(name="c0"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fp16_checkpoints_against_fp64(L, oracles, case):
    p = dict(CASES[case])
    w = synth.workload(p.pop("name"), **p)
    say(f"case: checkpoint precision {case} {w.nz}x{w.nx} nt={w.nt} shots={w.nshot}")
    obs = observed(oracles["f32s"], w)
    pre32, g32 = gpu_gradient(L, w, obs, 32)
    pre16, g16 = gpu_gradient(L, w, obs, 16)
    o64 = oracles["f64"]
    vp, rho = Array2D(w.vp_start), Array2D(w.rho_start)
    pre64 = obs.header_copy()
    g = Array2D(np.zeros((w.nx, w.nz), np.float32))
    o64.propagation(vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
                    w.freesurface, w.freq, float(w.filter), w.precondition, 1, pre64.recs, g.pp)
    g64 = g.a.copy()
    e = {"adjoint source fp32 store": rel(recs(pre32), recs(pre64)),
         "adjoint source fp16 store": rel(recs(pre16), recs(pre64)),
         "gradient fp32 store": rel(g32, g64),
         "gradient fp16 store": rel(g16, g64),
         "gradient fp16 vs fp32 store": rel(g16, g32)}
    say("  rel L2 against the FP64 oracle:", {k: f"{v:.3e}" for k, v in e.items()})
    assert np.array_equal(recs(pre16), recs(pre32))
    assert e["gradient fp32 store"] <= 1e-3
    assert e["gradient fp16 store"] <= 1e-2
    assert e["adjoint source fp32 store"] <= 1e-3


def test_fp16_checkpoints_same_in_both_layouts(L, oracles):
    """70 shots: FP16 slots in the shots-in-lanes layout (two 64-lane groups)
    and in per-shot arenas hold the same values, so the gradients agree bit
    for bit; and a batch of 33 shots gives the same bits again."""
    w = synth.workload("tiny", nshot=70, freesurface=1, shot_depth_cells=1, nx=90)
    say("case: fp16 checkpoints lanes vs arenas tiny fs=1 70 shots")
    obs = observed(oracles["f32s"], w)
    out = {}
    try:
        for name, layout, batch in (("arenas", 0, 0), ("lanes", 1, 0), ("lanes33", 1, 33)):
            L.fwiedu_set_layout(layout)
            L.fwiedu_set_batch(batch)
            out[name] = gpu_gradient(L, w, obs, 16)
    finally:
        L.fwiedu_set_layout(-1)
        L.fwiedu_set_batch(0)
    for name in ("lanes", "lanes33"):
        assert np.array_equal(out[name][1], out["arenas"][1]), name
        assert np.array_equal(recs(out[name][0]), recs(out["arenas"][0])), name
