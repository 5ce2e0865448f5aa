"""ctypes binding of libfluid_engine.so (include/fluid_engine.h).

The library is the product path: it is loaded from this package directory and every
call checks its status code, raising FluidEngineError with fe_last_error() on failure.
Nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int32, c_size_t,
                    c_uint32, c_uint64, c_void_p)

LIB_NAME = "libfluid_engine.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

FE_OK = 0
FE_ERR_UNSTABLE = 6
FE_HEALTH_NONFINITE = 1
FE_HEALTH_SPEED = 2
FE_SORT_BITONIC = 0
FE_SORT_RADIX = 1
FE_SOLVER_FIXED = 0
FE_SOLVER_CONVERGE = 1
FE_WARM_START_REFERENCE = 0
FE_WARM_START_LEARNED = 1
FE_DIVERGENCE_REFERENCE = 0
FE_DIVERGENCE_CONSISTENT = 1
FE_BOUNDARY_CLAMP = 0
FE_BOUNDARY_PARTICLES = 1
FE_RESIDUAL_COMPRESSION = 0
FE_RESIDUAL_TWO_SIDED = 1
FE_NET_DENSITY = 0
FE_NET_DIVERGENCE = 1
FE_MLP_INPUTS = 12
FE_MLP_HIDDEN = 64
FE_MLP_PARAMS = 9217
FE_MLP_NORM = 26

# field name -> (id, floats-or-words per particle, per-table (True) or per-particle)
FIELDS = {
    "position_a": (0, 4, False),
    "position_b": (1, 4, False),
    "velocity_a": (2, 4, False),
    "velocity_b": (3, 4, False),
    "colors": (4, 4, False),
    "densities": (5, 1, False),
    "factors": (6, 1, False),
    "source_terms": (7, 1, False),
    "pressures": (8, 1, False),
    "pressure_accelerations": (9, 4, False),
    "grid_entries": (10, 2, True),
    "grid_start": (11, 1, True),
}
INT_FIELDS = ("grid_entries", "grid_start")


class FeParams(Structure):
    """`SimulationParams` (src/entities/particle.rs:310-331), 112 bytes, std140."""

    _fields_ = [
        ("particle_radius", c_float),
        ("particle_mass", c_float),
        ("smoothing_radius", c_float),
        ("target_density", c_float),
        ("viscosity", c_float),
        ("relax_factor", c_float),
        ("dt", c_float),
        ("density_solver_iterations", c_uint32),
        ("divergence_solver_iterations", c_uint32),
        ("_pad", c_float * 3),
        ("gravity", c_float * 4),
        ("box_min", c_float * 4),
        ("box_max", c_float * 4),
        ("grid_res", c_int32 * 4),
    ]


class FeTuneEntry(Structure):
    _fields_ = [
        ("kernel", ctypes.c_char * 32),
        ("block", c_uint32),
        ("items", c_uint32),
        ("median_us", c_float),
        ("default_us", c_float),
    ]


class FeOptions(Structure):
    _fields_ = [
        ("sort_algorithm", c_uint32),
        ("use_cfl", c_uint32),
        ("use_solver_error_threshold", c_uint32),
        ("density_error_pct", c_float),
        ("divergence_error_pct", c_float),
        ("solver_mode", c_uint32),
        ("density_tol", c_float),
        ("divergence_tol", c_float),
        ("max_iterations", c_uint32),
        ("warm_start", c_uint32),
        ("divergence_order", c_uint32),
        ("boundary", c_uint32),
        ("residual", c_uint32),
        ("reserved", c_uint32 * 3),
    ]


class FeWall(Structure):
    """`CollisionBox` (src/entities/collision.rs:11-18) with its private time and base_min_x."""

    _fields_ = [
        ("min", c_float * 3),
        ("max", c_float * 3),
        ("wave_amplitude", c_float),
        ("wave_frequency", c_float),
        ("time", c_float),
        ("base_min_x", c_float),
    ]


class FeHealth(Structure):
    """fe_health: the watchdog's report of the last frame."""

    _fields_ = [
        ("frame", c_uint64),
        ("nonfinite", c_uint32),
        ("first_nonfinite", c_uint32),
        ("max_speed", c_float),
        ("speed_limit", c_float),
        ("flags", c_uint32),
        ("sort_error", c_uint32),
    ]


assert ctypes.sizeof(FeParams) == 112
assert ctypes.sizeof(FeWall) == 40
assert ctypes.sizeof(FeHealth) == 32


class FluidEngineError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where} failed (status {status}): {message}")
        self.status = status


_lib = None

# (name, argtypes)
_SIGNATURES = [
    ("fe_version", []),
    ("fe_last_error", []),
    ("fe_create", [POINTER(FeParams), c_void_p, c_uint32, c_int, POINTER(c_void_p)]),
    ("fe_destroy", [c_void_p]),
    ("fe_set_params", [c_void_p, POINTER(FeParams)]),
    ("fe_get_params", [c_void_p, POINTER(FeParams)]),
    ("fe_set_options", [c_void_p, POINTER(FeOptions)]),
    ("fe_get_options", [c_void_p, POINTER(FeOptions)]),
    ("fe_default_options", [POINTER(FeOptions)]),
    ("fe_num_particles", [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    ("fe_upload", [c_void_p, c_int, c_void_p, c_size_t]),
    ("fe_download", [c_void_p, c_int, c_void_p, c_size_t]),
    ("fe_device_ptr", [c_void_p, c_int, POINTER(c_void_p)]),
    ("fe_neighbor_search", [c_void_p]),
    ("fe_sort", [c_void_p]),
    ("fe_sort_pairs_noalloc", [c_int, c_void_p, c_uint32, c_void_p]),
    ("fe_density_alpha", [c_void_p]),
    ("fe_viscosity", [c_void_p]),
    ("fe_density_source", [c_void_p]),
    ("fe_pressure_force", [c_void_p]),
    ("fe_pressure_update", [c_void_p]),
    ("fe_pressure_integration", [c_void_p]),
    ("fe_divergence_source", [c_void_p]),
    ("fe_divergence_integration", [c_void_p]),
    ("fe_stats", [c_void_p]),
    ("fe_solve_density", [c_void_p, POINTER(c_uint32)]),
    ("fe_solve_divergence", [c_void_p, POINTER(c_uint32)]),
    ("fe_solve_report_enable", [c_void_p, c_int]),
    ("fe_solve_report", [c_void_p, c_int, POINTER(c_uint32), POINTER(c_double)]),
    ("fe_step", [c_void_p, c_float]),
    ("fe_run", [c_void_p, c_float, c_uint32]),
    ("fe_read_stats", [c_void_p, POINTER(c_uint32)]),
    ("fe_last_iterations", [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    ("fe_iteration_totals", [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    ("fe_reset_iteration_totals", [c_void_p]),
    ("fe_sync", [c_void_p]),
    ("fe_timer_start", [c_void_p]),
    ("fe_timer_stop", [c_void_p, POINTER(c_float)]),
    ("fe_profile_enable", [c_void_p, c_int]),
    ("fe_profile_count", [c_void_p, POINTER(c_uint32)]),
    ("fe_profile_entry", [c_void_p, c_uint32, c_char_p, c_size_t, POINTER(c_uint64), POINTER(c_double)]),
    ("fe_profile_reset", [c_void_p]),
    ("fe_frame_control", [POINTER(c_uint32), c_uint32, POINTER(FeParams), POINTER(FeOptions),
                          POINTER(c_float), POINTER(c_uint32), POINTER(c_uint32)]),
    ("fe_splat_density", [c_void_p]),
    ("fe_density_texture", [c_void_p, POINTER(c_void_p), POINTER(c_size_t)]),
    ("fe_download_density_texture", [c_void_p, c_void_p, c_size_t]),
    ("fe_copy_field_device", [c_void_p, c_int, c_void_p, c_size_t]),
    ("fe_wall_init", [POINTER(FeWall), POINTER(c_float), POINTER(c_float)]),
    ("fe_wall_update", [POINTER(FeWall), c_float]),
    ("fe_set_wall", [c_void_p, POINTER(FeWall)]),
    ("fe_get_wall", [c_void_p, POINTER(FeWall), POINTER(c_int)]),
    ("fe_learn_collect", [c_void_p, c_uint32, c_uint32]),
    ("fe_learn_dataset_size", [c_void_p, c_int, POINTER(c_uint32)]),
    ("fe_learn_dataset", [c_void_p, c_int, c_void_p, c_void_p, c_uint32]),
    ("fe_learn_train", [c_void_p, c_int, c_uint32, c_uint32, c_float, c_uint32, c_void_p]),
    ("fe_learn_get_weights", [c_void_p, c_int, c_void_p, c_void_p]),
    ("fe_learn_set_weights", [c_void_p, c_int, c_void_p, c_void_p]),
    ("fe_learn_forward", [c_void_p, c_int, c_void_p, c_uint32, c_void_p]),
    ("fe_learn_gradient", [c_void_p, c_int, c_void_p, c_void_p, c_uint32, c_void_p, POINTER(c_float)]),
    ("fe_learn_features", [c_void_p, c_int, c_void_p, c_void_p]),
    ("fe_learn_predict", [c_void_p, c_int]),
    ("fe_batch_create", [POINTER(FeParams), c_void_p, c_void_p, c_uint32, c_int, POINTER(c_void_p)]),
    ("fe_batch_scenes", [c_void_p, POINTER(c_uint32)]),
    ("fe_batch_scene_info", [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32)]),
    ("fe_batch_set_params", [c_void_p, c_uint32, POINTER(FeParams)]),
    ("fe_batch_download", [c_void_p, c_uint32, c_int, c_void_p, c_size_t]),
    ("fe_batch_upload", [c_void_p, c_uint32, c_int, c_void_p, c_size_t]),
    ("fe_batch_read_stats", [c_void_p, c_uint32, POINTER(c_uint32)]),
    ("fe_batch_iterations", [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint64),
                             POINTER(c_uint64)]),
    ("fe_batch_use_graph", [c_void_p, c_int]),
    ("fe_use_graph", [c_void_p, c_int]),
    ("fe_set_watchdog", [c_void_p, c_int, c_float, c_int]),
    ("fe_get_health", [c_void_p, POINTER(FeHealth)]),
    ("fe_batch_set_wall", [c_void_p, c_uint32, POINTER(FeWall)]),
    ("fe_batch_get_wall", [c_void_p, c_uint32, POINTER(FeWall), POINTER(c_int)]),
    ("fe_autotune", [c_void_p, c_uint32]),
    ("fe_tune_count", [c_void_p, POINTER(c_uint32)]),
    ("fe_tune_get", [c_void_p, c_uint32, POINTER(FeTuneEntry)]),
    ("fe_tune_set", [c_void_p, c_char_p, c_uint32, c_uint32]),
    ("fe_device_name", [c_void_p, c_char_p, c_size_t]),
    ("fe_check_exact_arith", [c_int, c_float, c_uint64, POINTER(c_uint64)]),
]

EXPORTED_SYMBOLS = tuple(name for name, _ in _SIGNATURES)


def load_library(path: str | None = None):
    """Load libfluid_engine.so; raise if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FluidEngineError(-1, "load_library",
                               f"{p} is missing; build it with `make -C fluid-engine_amd` "
                               "or __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    for name, argtypes in _SIGNATURES:
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_char_p if name in ("fe_version", "fe_last_error") else c_int32
    if path is None:
        _lib = lib
    return lib


def check(status: int, where: str) -> None:
    if status != FE_OK:
        msg = _lib.fe_last_error().decode() if _lib is not None else "?"
        raise FluidEngineError(status, where, msg)
