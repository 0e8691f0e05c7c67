"""ctypes binding of libfwiedu.so (the HIP library) with the reference's C signatures.

The library must be built first (``make -C fwi-edu_amd`` or
``__graft_entry__.build()``); loading fails loudly when it is missing, and every
compute entry point runs on the GPU -- there is no CPU fallback.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_long

from .records import Record2D, c_float_p, c_float_pp

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfwiedu.so")

_lib = None


class Acoustic2dCheckPoint(ctypes.Structure):
    """checkpoint.h:16-26"""
    _fields_ = [
        ("nt", c_int),
        ("dCheckPoint", c_int),
        ("nCheckPoint", c_int),
        ("ntPoint", c_int),
        ("status", POINTER(c_int)),
        ("memCost", c_float),
        ("p", POINTER(c_float_pp)),
        ("storage_p", POINTER(c_float_pp)),
    ]


class FwieduIteration(ctypes.Structure):
    """FwieduIteration of include/fwiedu.h (one FWI driver iteration)."""
    _fields_ = [("freq", c_float), ("filter", c_float), ("rz", c_int), ("rx", c_int),
                ("waterdepth", c_int), ("vp_water", c_float), ("rho_water", c_float),
                ("vpmin", c_float), ("vpmax", c_float), ("rhomin", c_float), ("rhomax", c_float),
                ("gardner", c_int), ("fraction", c_float),
                ("fun1", c_float), ("fun2", c_float), ("steplen", c_float), ("scalar", c_float),
                ("iter", c_int), ("verbose", c_int)]


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `make -C fwi-edu_amd` "
                           "(the FWI path runs only through the HIP library)")
    L = ctypes.CDLL(LIB_PATH)
    R = POINTER(Record2D)
    CK = POINTER(Acoustic2dCheckPoint)
    _sig(L, "iso_acoustic2d_propagation", None,
         [c_float_pp, c_float_pp, R, c_int, c_int, c_int, c_float, c_float, c_float, c_int, c_int,
          c_int, c_float, c_float, c_int, c_int, c_int, c_int, R, c_float_pp])
    _sig(L, "iso_acoustic2d_propagation_engine", None,
         [R, CK, c_float_pp, c_float_pp, c_int, c_int, c_float, c_float, c_float, c_int, c_int,
          c_int, c_int, c_int, c_int, c_int, c_float_pp])
    _sig(L, "acoustic_iso2d_fd_it_6order", None,
         [c_float_pp] * 10 + [c_float] * 4 + [c_int] * 4)
    _sig(L, "acoustic_iso2d_freesurface", None, [c_float_pp, c_int, c_int, c_int])
    _sig(L, "mealCoefs2d", None,
         [c_float_pp, c_float_pp, c_float_pp, c_float_pp, c_float, c_float, c_float, c_float,
          c_int, c_int, c_int])
    _sig(L, "acoustic2dCheckPoint_init", None, [CK])
    _sig(L, "alloc_acoustic2dCheckPoint", None, [CK, c_int, c_int, c_int])
    _sig(L, "free_acoustic2dCheckPoint", None, [CK])
    _sig(L, "save_acoustic2dCheckPoint", None, [CK, c_float_pp, c_float_pp, c_int, c_int])
    _sig(L, "get_acoustic2dCheckPoint", None, [CK, c_float_pp, c_float_pp, c_int, c_int])
    _sig(L, "get_acoustic2dWavefield", None, [CK, c_float_pp, c_int, c_int])
    _sig(L, "save_acoustic2dWavefield", None, [CK, c_float_pp, c_int, c_int])
    _sig(L, "get_acoustic2dCheckPointStatus", c_int, [CK, c_int])
    _sig(L, "calculate_memory_cost_checkpoint", c_float, [c_int, c_int, c_int])
    _sig(L, "compute_adjoint_source2d", None, [R, R, c_float_pp, c_float, c_int, R])
    _sig(L, "apply_gradient_precondition", None, [c_float_p, c_float_p, c_int, c_float])
    _sig(L, "apply_gradient_shot_precondition", None, [c_float_p, c_float_p, c_int])
    _sig(L, "calculate_objective_value", c_float, [R, c_int])
    _sig(L, "compute_model_perturbation", c_float, [c_float, c_float_p, c_float_p, c_float_p, c_int])
    _sig(L, "calculate_step_length", c_float, [R, R, c_int])
    _sig(L, "fwiedu_session_create", ctypes.c_void_p,
         [c_float_pp, c_float_pp, R, c_int, c_int, c_int, c_float, c_float, c_float, c_int, c_int,
          c_int, c_float, c_float, c_int])
    _sig(L, "fwiedu_session_set_model", None, [ctypes.c_void_p, c_float_pp, c_float_pp])
    _sig(L, "fwiedu_session_gradient", c_float, [ctypes.c_void_p, c_float_pp])
    _sig(L, "fwiedu_session_destroy", None, [ctypes.c_void_p])
    _sig(L, "fwiedu_session_iterate", None,
         [ctypes.c_void_p, POINTER(FwieduIteration), c_float_pp, c_float_pp, c_float_pp])
    _sig(L, "fwiedu_set_batch", None, [c_int])
    _sig(L, "fwiedu_set_tile", None, [c_int, c_int])
    _sig(L, "fwiedu_set_xcd_remap", None, [c_int])
    _sig(L, "fwiedu_set_march", None, [c_int, c_int, c_int])
    _sig(L, "fwiedu_set_tile_order", None, [c_int])
    _sig(L, "fwiedu_set_lanes", None, [c_int, c_int, c_int, c_int])
    _sig(L, "fwiedu_set_layout", None, [c_int])
    _sig(L, "fwiedu_set_graphs", None, [c_int])
    _sig(L, "fwiedu_set_inplace", None, [c_int])
    _sig(L, "fwiedu_set_fma", None, [c_int])
    _sig(L, "fwiedu_set_memory_fraction", None, [c_double])
    _sig(L, "fwiedu_set_kernel_timing", None, [c_int])
    _sig(L, "fwiedu_set_checkpoint_bits", None, [c_int])
    _sig(L, "calculate_peak_frequency", c_float, [c_float_p, c_int, c_float])
    _sig(L, "create_grid_observe2d", None, [R, c_int, c_float, c_float, c_int])
    _sig(L, "write_acoustic2d_traces", None, [c_char_p, R, c_int])
    _sig(L, "get_shots_number", c_int, [c_char_p])
    _sig(L, "read_gather2d_from_sufile", None, [R, c_char_p, c_int, c_int])
    _sig(L, "fwiedu_observe2d_from_json", None,
         [R, c_char_p, c_int, c_float_p, c_int, c_float_p, c_int, c_float, c_int, c_float_p, c_int])
    _sig(L, "fwiedu_json_shot_count", c_int, [c_char_p])
    _sig(L, "generate_shot_positions", c_int,
         [c_float_p, c_int, c_float_p, c_int, c_float, c_float, c_float, c_int,
          POINTER(c_float_p), POINTER(c_float_p)])
    _sig(L, "init_observe2d", None,
         [R, c_int, c_float_p, c_int, c_float_p, c_int, c_float, c_int, c_float_p, c_float_p,
          c_float_p, c_float, c_float, c_float, c_float, c_int])
    _sig(L, "fwiedu_stat", c_double, [c_char_p])
    _sig(L, "fwiedu_reset_stats", None, [])
    _sig(L, "fwiedu_lanes_max_shots", c_long, [c_int, c_int, c_int])
    _sig(L, "fwiedu_model_perturbation_noalloc", c_float,
         [c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_int])
    _sig(L, "fwiedu_release", None, [])
    _sig(L, "fwiedu_device_ok", c_int, [])
    _sig(L, "fwiedu_version", c_char_p, [])
    _lib = L
    return L


# Symbols of include/fwiedu.h, used by the load/export test.
EXPORTED = [
    "iso_acoustic2d_propagation", "iso_acoustic2d_propagation_engine",
    "acoustic_iso2d_fd_it_6order", "acoustic_iso2d_freesurface", "mealCoefs2d",
    "acoustic2dCheckPoint_init", "alloc_acoustic2dCheckPoint", "free_acoustic2dCheckPoint",
    "save_acoustic2dCheckPoint", "get_acoustic2dCheckPoint", "get_acoustic2dWavefield",
    "save_acoustic2dWavefield", "get_acoustic2dCheckPointStatus",
    "calculate_memory_cost_checkpoint", "compute_adjoint_source2d",
    "apply_gradient_precondition", "apply_gradient_shot_precondition",
    "calculate_objective_value", "compute_model_perturbation", "calculate_step_length",
    "fwiedu_session_create", "fwiedu_session_set_model", "fwiedu_session_gradient",
    "fwiedu_session_destroy", "fwiedu_session_iterate", "fwiedu_set_batch", "fwiedu_set_tile", "fwiedu_set_march", "fwiedu_set_tile_order", "fwiedu_set_xcd_remap", "fwiedu_set_lanes", "fwiedu_set_layout", "fwiedu_set_graphs", "fwiedu_set_inplace", "fwiedu_set_fma", "fwiedu_set_memory_fraction", "fwiedu_set_kernel_timing",
    "fwiedu_set_checkpoint_bits", "fwiedu_stat", "fwiedu_reset_stats", "fwiedu_release",
    "fwiedu_lanes_max_shots", "fwiedu_model_perturbation_noalloc",
    "fwiedu_device_ok", "fwiedu_version",
    "fwiedu_server_create", "fwiedu_server_submit", "fwiedu_server_done", "fwiedu_server_wait",
    "fwiedu_server_stat", "fwiedu_server_destroy",
    "calculate_peak_frequency", "create_grid_observe2d", "write_acoustic2d_traces",
    "get_shots_number", "read_gather2d_from_sufile", "fwiedu_observe2d_from_json",
    "generate_shot_positions", "init_observe2d",
    "fwiedu_json_shot_count",
]


def stat(name):
    return load().fwiedu_stat(name.encode())


def propagate(lib, w, vp, rho, obs, problem, pre=None, grad=None, precondition=None, iter=1,
              verbose=0):
    """iso_acoustic2d_propagation with a synth.Workload's parameters.
    vp/rho are records.Array2D; obs/pre are records.ShotSet."""
    pc = w.precondition if precondition is None else precondition
    lib.iso_acoustic2d_propagation(
        vp.pp, rho.pp, obs.recs, w.nshot, w.nz, w.nx, w.dx, w.dz, w.dt, w.nt, w.pml,
        w.freesurface, w.freq, float(w.filter), pc, problem, iter, verbose,
        (pre or obs).recs, grad.pp if grad is not None else None)
