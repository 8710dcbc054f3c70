"""ctypes binding of libnbody_sim.so (declared in include/nbody_sim.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C webgpu-fluid-sim_amd/csrc``).  Loading fails loudly when it is
missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbody_sim.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_HIP = -2
ERR_CAPACITY = -3
ERR_STACK_OVERFLOW = -4
ERR_STATE = -5
ERR_NO_DEVICE = -6

SOFTEN_REFERENCE = 0
SOFTEN_CONSTANT = 1
MORTON_REFERENCE = 0
MORTON_BBOX30 = 1
MATH_FAST = 0
MATH_EXACT = 1
EXPANSION_MONOPOLE = 0
EXPANSION_QUADRUPOLE = 1

DEBUG_ACCEL = 1
DEBUG_STATS = 2
DEBUG_FILL_CLIMB = 4
DEBUG_SORT_RADIX = 8

BUF_KEYS = 1
BUF_BODY_INDICES = 2
BUF_CHILDREN = 3
BUF_PARENT = 4
BUF_NODE_CM = 5
BUF_NODE_AABB = 6
BUF_ACCEL = 7
BUF_BBOX = 8
BUF_SIM_ROOTS = 9
BUF_WALK_COUNTS = 10
BUF_TONE_TABLE = 11
BUF_NODE_QUAD = 12

SCENARIO_DEFAULT = 0
SCENARIO_TWO_CLUSTERS = 1
SCENARIO_PLANETARY_DISK = 2
SCENARIO_UNIFORM_DISK = 3
SCENARIO_EXP_DISK = 4
SCENARIO_COLLISION = 5
SCENARIO_PLUMMER = 6

PHASES = ("bbox", "encode", "sort", "build", "fill", "walk")

# Every symbol include/nbody_sim.h declares.
EXPORTS = (
    "nbody_default_config", "nbody_create", "nbody_ensemble_create", "nbody_destroy",
    "nbody_set_config", "nbody_get_config", "nbody_set_bodies", "nbody_get_bodies",
    "nbody_get_num_bodies", "nbody_step", "nbody_synchronize", "nbody_set_stream",
    "nbody_use_graph", "nbody_prepare_graph", "nbody_set_debug", "nbody_debug_get", "nbody_walk_stats",
    "nbody_set_profiling", "nbody_phase_times", "nbody_set_tuning", "nbody_get_tuning",
    "nbody_last_error",
    "nbody_generate", "nbody_philox4x32", "nbody_version", "nbody_adjoint_forward",
    "nbody_adjoint_backward", "nbody_step_static", "nbody_direct_accel_f64",
    "nbody_set_radius_multiplier", "nbody_get_radius_multiplier", "nbody_get_step_count",
    "nbody_save_snapshot", "nbody_load_snapshot", "nbody_generate_radius",
    "nbody_default_camera", "nbody_render", "nbody_write_png",
)

# Largest body count of nbody_step_static (its fixed device scratch).
STATIC_MAX_BODIES = 65536


class NbodyConfig(ctypes.Structure):
    _fields_ = [
        ("grav_constant", ctypes.c_float),
        ("bh_theta", ctypes.c_float),
        ("substeps", ctypes.c_int32),
        ("delta_time", ctypes.c_float),
        ("epsilon_multiplier", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("softening_mode", ctypes.c_int32),
        ("morton_mode", ctypes.c_int32),
        ("math_mode", ctypes.c_int32),
        ("offset", ctypes.c_uint32),
        ("expansion", ctypes.c_int32),
    ]


class NbodyTuning(ctypes.Structure):
    _fields_ = [
        ("walk_block", ctypes.c_int32),
        ("walk_stack", ctypes.c_int32),
        ("walk_variant", ctypes.c_int32),
        ("xcd_order", ctypes.c_int32),
        ("lane_bodies", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32),
        ("wave_bodies", ctypes.c_int32),
    ]


class NbodyCamera(ctypes.Structure):
    _fields_ = [
        ("center_x", ctypes.c_float),
        ("center_y", ctypes.c_float),
        ("half_x", ctypes.c_float),
        ("half_y", ctypes.c_float),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
    ]


class NbodyIcParams(ctypes.Structure):
    _fields_ = [
        ("grav_constant", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("dispersion", ctypes.c_float),
        ("mass_sigma", ctypes.c_float),
    ]


class NbodyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nbody error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load libnbody_sim.so once and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libnbody_sim.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32
    cfgp = ctypes.POINTER(NbodyConfig)
    sig = {
        "nbody_default_config": (ctypes.c_int, [cfgp]),
        "nbody_create": (vp, [cfgp, i64, i32, ctypes.POINTER(ctypes.c_int)]),
        "nbody_ensemble_create": (vp, [cfgp, i32, ctypes.POINTER(ctypes.c_int64), i32,
                                       ctypes.POINTER(ctypes.c_int)]),
        "nbody_destroy": (ctypes.c_int, [vp]),
        "nbody_set_config": (ctypes.c_int, [vp, cfgp]),
        "nbody_get_config": (ctypes.c_int, [vp, cfgp]),
        "nbody_set_bodies": (ctypes.c_int, [vp, i64, vp, vp, vp]),
        "nbody_get_bodies": (ctypes.c_int, [vp, vp, vp, vp]),
        "nbody_get_num_bodies": (ctypes.c_int64, [vp]),
        "nbody_step": (ctypes.c_int, [vp, i32]),
        "nbody_synchronize": (ctypes.c_int, [vp]),
        "nbody_set_stream": (ctypes.c_int, [vp, vp]),
        "nbody_use_graph": (ctypes.c_int, [vp, i32]),
        "nbody_prepare_graph": (ctypes.c_int, [vp]),
        "nbody_set_debug": (ctypes.c_int, [vp, i32]),
        "nbody_debug_get": (ctypes.c_int, [vp, i32, vp, i64]),
        "nbody_walk_stats": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64), i32]),
        "nbody_set_profiling": (ctypes.c_int, [vp, i32]),
        "nbody_phase_times": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(ctypes.c_int64), i32]),
        "nbody_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(NbodyTuning)]),
        "nbody_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(NbodyTuning)]),
        "nbody_last_error": (ctypes.c_char_p, [vp]),
        "nbody_generate": (ctypes.c_int, [i32, i64, ctypes.c_uint64,
                                          ctypes.POINTER(NbodyIcParams), vp, vp, vp]),
        "nbody_philox4x32": (None, [ctypes.POINTER(u32), ctypes.POINTER(u32),
                                    ctypes.POINTER(u32)]),
        "nbody_version": (ctypes.c_char_p, []),
        "nbody_adjoint_forward": (ctypes.c_int, [vp, i32, i32]),
        "nbody_adjoint_backward": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "nbody_step_static": (ctypes.c_int, [cfgp, i64, vp, vp, vp, i32, vp]),
        "nbody_direct_accel_f64": (ctypes.c_int, [vp, vp, i64, vp]),
        "nbody_set_radius_multiplier": (ctypes.c_int, [vp, vp]),
        "nbody_get_radius_multiplier": (ctypes.c_int, [vp, vp]),
        "nbody_get_step_count": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "nbody_save_snapshot": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "nbody_load_snapshot": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "nbody_generate_radius": (ctypes.c_int, [i32, i64, vp]),
        "nbody_default_camera": (ctypes.c_int, [i32, i32, ctypes.POINTER(NbodyCamera)]),
        "nbody_render": (ctypes.c_int, [vp, ctypes.POINTER(NbodyCamera), vp, vp]),
        "nbody_write_png": (ctypes.c_int, [ctypes.c_char_p, i32, i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error(handle=None) -> str:
    msg = lib().nbody_last_error(handle)
    return msg.decode() if msg else ""


def check(rc: int, handle=None) -> None:
    if rc != OK:
        raise NbodyError(rc, last_error(handle))
