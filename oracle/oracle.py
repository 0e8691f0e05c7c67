"""ctypes loader for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to check
(or time beside) the HIP library.  The library itself never imports this.
"""
import ctypes
import os
from ctypes import POINTER, c_float, c_int, c_long

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
c_float_p = POINTER(c_float)
c_float_pp = POINTER(c_float_p)
c_int_p = POINTER(c_int)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Oracle:
    def __init__(self, precision="f32", path=None):
        path = path or os.path.join(HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        from fwiedu.records import Record2D  # same struct layout as OracleRecord2D
        R = POINTER(Record2D)
        prop_args = [c_float_pp, c_float_pp, R, c_int, c_int, c_int, c_float, c_float, c_float,
                     c_int, c_int, c_int, c_float, c_float, c_int, c_int, R, c_float_pp]
        self.propagation = _sig(L, "oracle_propagation", None, prop_args)
        self.propagation_verbose = _sig(L, "oracle_propagation_verbose", None,
                                        prop_args + [c_int, c_int])
        self.gradient_fullstore = _sig(L, "oracle_gradient_fullstore", None,
                                       prop_args[:15] + [R, c_float_pp])
        self.checkpoint_plan_ = _sig(L, "oracle_checkpoint_plan", None,
                                     [c_int, c_int_p, c_int_p, c_int_p, c_int_p])
        self.memory_cost_checkpoint = _sig(L, "oracle_memory_cost_checkpoint", c_float,
                                           [c_int, c_int, c_int])
        self.fd_step = _sig(L, "oracle_fd_step", None,
                            [c_float_p] * 10 + [c_int, c_int, c_float, c_float, c_float, c_float,
                                                c_int, c_int, c_int, c_int])
        self.freesurface = _sig(L, "oracle_freesurface", None, [c_float_p, c_int, c_int, c_int])
        self.setup_coefs = _sig(L, "oracle_setup_coefs", None,
                                [c_float_p, c_float_p, c_int, c_int, c_int, c_float, c_float,
                                 c_float, c_float] + [c_float_p] * 9)
        self.meal_coefs = _sig(L, "oracle_meal_coefs", None,
                               [c_float_p, c_float_p, c_float_p, c_float, c_float, c_float, c_int,
                                c_int, c_int])
        self.pad_model = _sig(L, "oracle_pad_model", None, [c_float_p, c_int, c_int, c_float_p, c_int])
        self.adjoint_source = _sig(L, "oracle_adjoint_source", None,
                                   [R, R, c_float_pp, c_float, c_int, R])
        self.objective = _sig(L, "oracle_objective", c_float, [R, c_int])
        self.step_length = _sig(L, "oracle_step_length", c_float, [R, R, c_int])
        self.model_perturbation = _sig(L, "oracle_model_perturbation", c_float,
                                       [c_float, c_float_p, c_float_p, c_float_p, c_int])
        self.shot_precondition = _sig(L, "oracle_shot_precondition", None, [c_float_p, c_float_p, c_int])
        self.global_precondition = _sig(L, "oracle_global_precondition", None,
                                        [c_float_p, c_float_p, c_int, c_float])
        self.taper_traces = _sig(L, "oracle_taper_traces", None, [c_float_pp, c_int, c_int, c_int, c_int])
        self.smooth2d = _sig(L, "oracle_smooth2d", None, [c_float_pp, c_float_pp, c_int, c_int, c_int, c_int])
        self.peak_frequency = _sig(L, "oracle_peak_frequency", c_float, [c_float_p, c_int, c_float])
        self.vp2density = _sig(L, "oracle_vp2density", None, [c_float_p, c_float_p, c_int])
        self.clip = _sig(L, "oracle_clip", None, [c_float_p, c_int, c_float, c_float])
        self.v2slow = _sig(L, "oracle_v2slow", None, [c_float_p, c_float_p, c_long])
        self.slow2v = _sig(L, "oracle_slow2v", None, [c_float_p, c_float_p, c_long])
        self.set_water = _sig(L, "oracle_set_water", None, [c_float_pp, c_float, c_int, c_int, c_int])
        self.lowpass_traces = _sig(L, "oracle_lowpass_traces", None,
                                   [c_float_pp, c_float_pp, c_int, c_int, c_float, c_float, c_float])
        self.bandpass_traces = _sig(L, "oracle_bandpass_traces", None,
                                    [c_float_pp, c_float_pp, c_int, c_int, c_float, c_float, c_float])
        self.interpolate_traces = _sig(L, "oracle_interpolate_traces", None,
                                       [c_float_pp, c_int, c_int, c_int, c_float_pp])
        self.is_f64 = bool(_sig(L, "oracle_is_f64", c_int, [])())

    def checkpoint_plan(self, nt):
        a, b, c, d = c_int(), c_int(), c_int(), c_int()
        self.checkpoint_plan_(nt, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
        return a.value, b.value, c.value, d.value
