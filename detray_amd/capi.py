"""ctypes mirror of the C ABI declared in include/detray_amd.h.

The structures here are byte-for-byte the C structs; the shared library
``libdetray_amd.so`` (built in-tree by ``detray_amd.build``) is loaded lazily
and the product path raises if it is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libdetray_amd.so"

DTR_SUCCESS = 0
STATUS_NAMES = {
    0: "DTR_SUCCESS",
    1: "DTR_ERROR_INVALID_ARGUMENT",
    2: "DTR_ERROR_HIP",
    3: "DTR_ERROR_OUT_OF_MEMORY",
    4: "DTR_ERROR_NO_DEVICE",
    5: "DTR_ERROR_UNSUPPORTED",
    6: "DTR_ERROR_IO",
}

# dtr_track_status
TRACK_EXIT, TRACK_ABORT, TRACK_PATH_LIMIT, TRACK_OVERFLOW, TRACK_MOMENTUM_LIMIT, TRACK_TARGET = 0, 1, 2, 3, 4, 5
TRACK_PAUSED = 6
# dtr_mask_shape
RECTANGLE2D, TRAPEZOID2D, CONCENTRIC_CYLINDER2D, RING2D, LINE_SQUARE, LINE_CIRCULAR = 0, 1, 2, 3, 4, 5
ANNULUS2D, CYLINDER2D = 6, 7
MATERIAL_SLAB, MATERIAL_ROD = 0, 1
# dtr_surface_id
PORTAL, SENSITIVE, PASSIVE = 0, 1, 2
INVALID_INDEX = 0xFFFFFFFF
END_OF_WORLD = 0xFFFF


class Transform(C.Structure):
    _fields_ = [("r", C.c_double * 9), ("t", C.c_double * 3)]


class Mask(C.Structure):
    _fields_ = [
        ("v", C.c_double * 7),
        ("volume_link", C.c_uint16),
        ("shape", C.c_uint8),
        ("pad", C.c_uint8 * 5),
    ]


class MaterialSlab(C.Structure):
    _fields_ = [
        (n, C.c_double)
        for n in ("X0", "L0", "A", "Z", "mass_rho", "molar_rho", "mean_exc",
                  "thickness", "thickness_in_X0")
    ] + [("kind", C.c_uint32), ("has_density", C.c_uint32), ("density", C.c_double * 5)]


class Surface(C.Structure):
    _fields_ = [
        ("barcode", C.c_uint64),
        ("transform", C.c_uint32),
        ("mask", C.c_uint32),
        ("material", C.c_uint32),
        ("volume", C.c_uint16),
        ("n_masks", C.c_uint8),
        ("sf_id", C.c_uint8),
        ("shape", C.c_uint8),
        ("pad", C.c_uint8 * 7),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("min0", C.c_double), ("max0", C.c_double),
        ("min1", C.c_double), ("max1", C.c_double),
        ("radius", C.c_double),
        ("n0", C.c_uint32), ("n1", C.c_uint32),
        ("bins", C.c_uint32),
        ("type", C.c_uint8),
        ("capacity", C.c_uint8),
        ("pad", C.c_uint8 * 2),
    ]


class Volume(C.Structure):
    _fields_ = [
        ("transform", C.c_uint32),
        ("bf_begin", C.c_uint32), ("bf_end", C.c_uint32),
        ("sf_begin", C.c_uint32), ("sf_end", C.c_uint32),
        ("grid", C.c_int32),
        ("material", C.c_uint32),
    ]


MATERIAL_MAP = 0x80000000


class MaterialMap(C.Structure):
    _fields_ = [
        ("n0", C.c_uint32), ("n1", C.c_uint32),
        ("slabs", C.c_uint32),
        ("circular0", C.c_uint8), ("circular1", C.c_uint8), ("pad", C.c_uint8 * 2),
        ("min0", C.c_double), ("max0", C.c_double),
        ("min1", C.c_double), ("max1", C.c_double),
    ]


class DetectorDesc(C.Structure):
    _fields_ = [
        ("n_volumes", C.c_uint32), ("n_surfaces", C.c_uint32),
        ("n_transforms", C.c_uint32), ("n_masks", C.c_uint32),
        ("n_materials", C.c_uint32), ("n_grids", C.c_uint32),
        ("n_bf", C.c_uint32), ("n_bins", C.c_uint32),
        ("volumes", C.POINTER(Volume)),
        ("surfaces", C.POINTER(Surface)),
        ("transforms", C.POINTER(Transform)),
        ("masks", C.POINTER(Mask)),
        ("materials", C.POINTER(MaterialSlab)),
        ("grids", C.POINTER(Grid)),
        ("bf_surfaces", C.POINTER(C.c_uint32)),
        ("grid_bins", C.POINTER(C.c_uint32)),
        ("n_material_maps", C.c_uint32), ("pad", C.c_uint32),
        ("material_maps", C.POINTER(MaterialMap)),
    ]


class FieldDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("pad", C.c_int32),
        ("b", C.c_double * 3),
        ("n", C.c_uint32 * 3), ("pad2", C.c_uint32),
        ("min", C.c_double * 3),
        ("pitch", C.c_double * 3),
        ("values", C.POINTER(C.c_float)),
    ]


class PropConfig(C.Structure):
    _fields_ = [
        ("min_stepsize", C.c_float),
        ("rk_error_tol", C.c_float),
        ("step_constraint", C.c_float),
        ("path_limit", C.c_float),
        ("max_rk_updates", C.c_uint32),
        ("use_mean_loss", C.c_uint8),
        ("use_eloss_gradient", C.c_uint8),
        ("use_field_gradient", C.c_uint8),
        ("do_covariance_transport", C.c_uint8),
        ("min_mask_tolerance", C.c_float),
        ("max_mask_tolerance", C.c_float),
        ("mask_tolerance_scalor", C.c_float),
        ("path_tolerance", C.c_float),
        ("overstep_tolerance", C.c_float),
        ("search_window", C.c_uint32 * 2),
        ("accumulated_error", C.c_float),
        ("n_scattering_stddev", C.c_int32),
        ("estimate_scattering_noise", C.c_uint8),
        ("do_energy_loss", C.c_uint8),
        ("do_multiple_scattering", C.c_uint8),
        ("use_material", C.c_uint8),
        ("max_surfaces", C.c_uint32),
        ("pad", C.c_uint32),
        ("actors", C.c_void_p),
    ]


ACTOR_MOMENTUM_ABORTER, ACTOR_TARGET_ABORTER, ACTOR_STEP_TRACER = 1, 2, 4


class StepTrace(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32), ("pad", C.c_uint32),
        ("n", C.c_void_p), ("surface", C.c_void_p), ("step_size", C.c_void_p),
        ("path_length", C.c_void_p), ("n_trials", C.c_void_p), ("free_params", C.c_void_p),
        ("bound_params", C.c_void_p), ("jacobian", C.c_void_p),
    ]


class ActorStates(C.Structure):
    _fields_ = [
        ("actors", C.c_uint32), ("min_p", C.c_float), ("min_pT", C.c_float), ("pad", C.c_uint32),
        ("target_surface", C.c_void_p), ("trace", StepTrace),
    ]


class TrackBatch(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("pad", C.c_uint32),
        ("free_params", C.c_void_p),
        ("initial_cov", C.c_void_p),
        ("bound_params", C.c_void_p),
        ("start_surface", C.c_void_p),
    ]


class ResultBatch(C.Structure):
    _fields_ = [
        ("bound_params", C.c_void_p),
        ("bound_cov", C.c_void_p),
        ("free_params", C.c_void_p),
        ("path_length", C.c_void_p),
        ("surfaces", C.c_void_p),
        ("n_surfaces", C.c_void_p),
        ("n_steps", C.c_void_p),
        ("n_trials", C.c_void_p),
        ("status", C.c_void_p),
    ]


# dtr_precision, dtr_schedule
PRECISION_FP64, PRECISION_MIXED = 0, 1
SCHEDULE_GRID, SCHEDULE_PERSISTENT = 0, 1
NAVIGATION_WAVE, NAVIGATION_LANE, NAVIGATION_GATHER = 0, 1, 2
# dtr_nav_direction
DTR_DIRECTION_FORWARD, DTR_DIRECTION_BACKWARD = 0, 1


class LaunchConfig(C.Structure):
    _fields_ = [
        ("precision", C.c_uint32),
        ("schedule", C.c_uint32),
        ("block_size", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
        ("tracks_per_claim", C.c_uint32),
        ("min_idle", C.c_uint32),
        ("navigation", C.c_uint32),
        ("xcd_chunk", C.c_uint32),
        ("order", C.c_void_p),
        ("queue", C.c_void_p),
    ]


class SortConfig(C.Structure):
    _fields_ = [
        ("eta_bins_per_unit", C.c_uint32),
        ("phi_bins", C.c_uint32),
        ("eta_max", C.c_float),
        ("pad", C.c_uint32),
    ]


# dtr_adjoint_model, dtr_tape_status
ADJOINT_EXACT, ADJOINT_TRANSPORT = 0, 1
TAPE_OK, TAPE_STEP_OVERFLOW, TAPE_EVENT_OVERFLOW, TAPE_UNSUPPORTED = 0, 1, 2, 3


class AdjointConfig(C.Structure):
    _fields_ = [
        ("max_steps", C.c_uint32),
        ("max_events", C.c_uint32),
        ("checkpoint_every", C.c_uint32),
        ("precision", C.c_uint32),
        ("model", C.c_uint32),
        ("block_size", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
        ("pad", C.c_uint32),
        ("sigma_loc0", C.c_double),
        ("sigma_loc1", C.c_double),
    ]


class Hits(C.Structure):
    _fields_ = [
        ("max_hits", C.c_uint32),
        ("pad", C.c_uint32),
        ("surface", C.c_void_p),
        ("loc", C.c_void_p),
        ("n_hits", C.c_void_p),
    ]


class AdjointBatch(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("pad", C.c_uint32),
        ("bound_params", C.c_void_p),
        ("start_surface", C.c_void_p),
        ("tape", C.c_void_p),
        ("tape_bytes", C.c_size_t),
        ("order", C.c_void_p),
    ]


class SurfaceSequence(C.Structure):
    _fields_ = [
        ("max_surfaces", C.c_uint32),
        ("direction", C.c_uint32),
        ("surfaces", C.c_void_p),
        ("n_surfaces", C.c_void_p),
    ]


class PropagationState(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("max_steps", C.c_uint32),
        ("resume", C.c_uint32),
    ]


class WireChamberConfig(C.Structure):
    _fields_ = [
        ("n_layers", C.c_uint32),
        ("square_cells", C.c_uint32),
        ("first_layer_inner_radius", C.c_double),
        ("half_z", C.c_double),
        ("cell_size", C.c_double),
        ("stereo_angle", C.c_double),
        ("mat_radius", C.c_double),
    ]


class ToyConfig(C.Structure):
    _fields_ = [
        ("n_brl_layers", C.c_uint32),
        ("n_edc_layers", C.c_uint32),
        ("use_material", C.c_uint8),
        ("pad", C.c_uint8 * 3),
    ]


class TelescopeConfig(C.Structure):
    _fields_ = [
        ("n_planes", C.c_uint32), ("pad", C.c_uint32),
        ("positions", C.POINTER(C.c_double)),
        ("half_x", C.c_double), ("half_y", C.c_double),
        ("thickness", C.c_double),
        ("envelope", C.c_double),
        ("volume_material", C.POINTER(MaterialSlab)),
    ]


class TrackGenConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n", C.c_uint32),
        ("kind", C.c_uint32),
        ("p_min", C.c_double), ("p_max", C.c_double), ("sigma_theta", C.c_double),
        ("eta_min", C.c_double), ("eta_max", C.c_double),
        ("pt_min", C.c_double), ("pt_max", C.c_double),
        ("sigma_xy", C.c_double), ("sigma_z", C.c_double),
    ]


# Every symbol the header declares, with its ctypes signature.
_SIGNATURES = {
    "dtr_default_config": (None, [C.POINTER(PropConfig)]),
    "dtr_detector_create": (C.c_int, [C.POINTER(DetectorDesc), C.POINTER(C.c_void_p)]),
    "dtr_detector_destroy": (C.c_int, [C.c_void_p]),
    "dtr_field_create": (C.c_int, [C.POINTER(FieldDesc), C.POINTER(C.c_void_p)]),
    "dtr_field_destroy": (C.c_int, [C.c_void_p]),
    "dtr_propagate": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig),
                                C.POINTER(TrackBatch), C.POINTER(ResultBatch), C.c_void_p]),
    "dtr_default_launch": (None, [C.POINTER(LaunchConfig)]),
    "dtr_default_actors": (None, [C.POINTER(ActorStates)]),
    "dtr_propagate_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig),
                                   C.POINTER(TrackBatch), C.POINTER(ResultBatch),
                                   C.POINTER(LaunchConfig), C.c_void_p]),
    "dtr_default_sort": (None, [C.POINTER(SortConfig)]),
    "dtr_sort_workspace_size": (C.c_size_t, [C.POINTER(SortConfig)]),
    "dtr_sort_tracks": (C.c_int, [C.POINTER(SortConfig), C.c_void_p, C.c_uint32, C.c_void_p,
                                  C.c_void_p, C.c_size_t, C.c_void_p]),
    "dtr_sort_tracks_static": (C.c_int, [C.POINTER(SortConfig), C.c_void_p, C.c_uint32, C.c_void_p,
                                         C.c_void_p]),
    "dtr_propagate_direct": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig), C.POINTER(TrackBatch),
                                       C.POINTER(SurfaceSequence), C.POINTER(ResultBatch),
                                       C.POINTER(LaunchConfig), C.c_void_p]),
    "dtr_propagation_state_size": (C.c_size_t, [C.c_uint32]),
    "dtr_propagate_resumable": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig), C.POINTER(TrackBatch),
                                          C.POINTER(ResultBatch), C.POINTER(LaunchConfig),
                                          C.POINTER(PropagationState), C.c_void_p]),
    "dtr_default_adjoint": (None, [C.POINTER(AdjointConfig)]),
    "dtr_adjoint_tape_size": (C.c_size_t, [C.POINTER(AdjointConfig), C.c_uint32]),
    "dtr_adjoint_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig), C.POINTER(AdjointConfig),
                                      C.POINTER(AdjointBatch), C.POINTER(Hits), C.POINTER(Hits), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    "dtr_adjoint_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PropConfig), C.POINTER(AdjointConfig),
                                       C.POINTER(AdjointBatch), C.POINTER(Hits), C.c_void_p, C.c_void_p]),
    "dtr_build_toy_detector": (C.c_int, [C.POINTER(ToyConfig), C.POINTER(DetectorDesc)]),
    "dtr_build_wire_chamber": (C.c_int, [C.POINTER(WireChamberConfig), C.POINTER(DetectorDesc)]),
    "dtr_build_telescope_detector": (C.c_int, [C.POINTER(TelescopeConfig), C.POINTER(DetectorDesc)]),
    "dtr_predefined_material": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(MaterialSlab)]),
    "dtr_host_detector_free": (None, [C.POINTER(DetectorDesc)]),
    "dtr_check_detector": (C.c_int, [C.POINTER(DetectorDesc), C.c_char_p, C.c_size_t]),
    "dtr_read_detector_json": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                         C.POINTER(DetectorDesc), C.c_char_p, C.c_size_t]),
    # the view records are filled by the C++ adapter (include/detray_amd.hpp)
    "dtr_flatten_detector": (C.c_int, [C.c_void_p, C.POINTER(DetectorDesc), C.c_char_p, C.c_size_t]),
    "dtr_write_detector_json": (C.c_int, [C.POINTER(DetectorDesc), C.c_char_p, C.POINTER(C.c_char_p),
                                          C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                          C.c_size_t]),
    "dtr_sample_solenoid": (C.c_int, [C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "dtr_generate_tracks": (C.c_int, [C.POINTER(TrackGenConfig), C.POINTER(C.c_double)]),
    "dtr_version": (C.c_char_p, []),
}

HEADER_SYMBOLS = tuple(_SIGNATURES)

_lib = None


class DetrayError(RuntimeError):
    """Raised for any non-zero status returned by the C ABI."""


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libdetray_amd.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise DetrayError(
            f"{p} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int, what: str) -> None:
    if code != DTR_SUCCESS:
        raise DetrayError(f"{what} failed: {STATUS_NAMES.get(code, code)}")
