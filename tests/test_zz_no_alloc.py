"""fwiedu_model_perturbation_noalloc (device arrays, scratch in a fixed device
array, no device allocation) against compute_model_perturbation (host arrays)
on the same values: the same scalar and the same output bit for bit, on two
calls in a row, the second with a smaller input (the scratch keeps the first
call's contents; the histogram counters must be cleared by the call)."""
import ctypes

import numpy as np
import pytest

from fwiedu import lib as flib
from fwiedu.records import fptr

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs to hold device arrays."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        return p

    def upload(self, p, a):
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, a, p):
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2) == 0

    def free(self, p):
        self.h.hipFree(p)


@pytest.fixture(scope="module")
def L():
    lib = flib.load()
    assert lib.fwiedu_device_ok() == 1, "no HIP device"
    return lib


def case(n, seed):
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal(n) * 2.5 + 0.3).astype(np.float32)
    slow = rng.uniform(1.0 / 5000, 1.0 / 1500, n).astype(np.float32)
    return g, slow


def test_noalloc_perturbation_matches_existing_entry_point(L):
    hip = Hip()
    ptrs = []
    try:
        for n, seed in ((300000, 1), (4097, 2)):
            print(f"case: model perturbation n={n}", flush=True)
            g, slow = case(n, seed)
            # the existing entry point, host arrays
            out_h = np.zeros_like(g)
            sc_h = L.compute_model_perturbation(0.01, fptr(slow.copy()), fptr(g.copy()),
                                                fptr(out_h), n)
            # the no-allocation variant, device arrays the caller holds
            dg, ds, do = hip.malloc(g.nbytes), hip.malloc(slow.nbytes), hip.malloc(g.nbytes)
            ptrs += [dg, ds, do]
            hip.upload(dg, g)
            hip.upload(ds, slow)
            sc_d = L.fwiedu_model_perturbation_noalloc(0.01, ds, dg, do, n)
            out_d = np.zeros_like(g)
            hip.download(out_d, do)
            assert sc_d == sc_h, (n, sc_d, sc_h)
            assert np.array_equal(out_d, out_h), n
            assert np.abs(out_h).max() > 0
    finally:
        for p in ptrs:
            hip.free(p)
