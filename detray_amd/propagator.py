"""Host-side mirror of detray's propagation interface over the C ABI.

The names follow the reference:

* :class:`PropagationConfig` mirrors ``propagation::config``
  (core/include/detray/propagator/propagation_config.hpp: stepping,
  navigation and intersection configs with their defaults).
* :class:`Detector` holds the flattened ``detector<metadata>`` tables
  (toy or telescope geometry) and its device copy
  (``detector::view_type``, core/include/detray/core/detector.hpp:136-202).
* :class:`MagneticField` is the field view (constant or sampled grid).
* :class:`Propagator` runs ``propagator<rk_stepper, caching_navigator,
  actor_chain<pathlimit_aborter, parameter_updater<
  pointwise_material_interactor>, surface_sequencer>>::propagate`` over a
  batch of tracks on the GPU.

PyTorch is only used for device memory and the stream; the compute is the
hand-written HIP kernel in ``libdetray_amd.so``.
"""

from __future__ import annotations

import ctypes as C
import os
import math
from dataclasses import dataclass, field as dc_field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from . import capi

# detray units (core/include/detray/definitions/units.hpp)
MM = 1.0
M = 1000.0
UM = 1e-3
GEV = 1.0
MEV = 1e-3
TESLA = 0.000299792458


def _f32(x: float) -> float:
    return float(np.float32(x))


@dataclass
class PropagationConfig:
    """``propagation::config`` with the reference defaults."""

    # stepping::config (stepping_config.hpp:27-48)
    min_stepsize: float = 1e-4 * MM
    rk_error_tol: float = 1e-4 * MM
    step_constraint: float = float(np.finfo(np.float32).max)
    path_limit: float = 5.0 * M
    max_rk_updates: int = 10000
    use_mean_loss: bool = True
    use_eloss_gradient: bool = False
    use_field_gradient: bool = False
    do_covariance_transport: bool = True
    # intersection::config (intersection_config.hpp:20-36)
    min_mask_tolerance: float = 1e-5 * MM
    max_mask_tolerance: float = 3.0 * MM
    mask_tolerance_scalor: float = 5e-2
    path_tolerance: float = 1.0 * UM
    overstep_tolerance: float = -1000.0 * UM
    # navigation::config (navigation_config.hpp:23-40)
    search_window: Sequence[int] = (0, 0)
    accumulated_error: float = 0.001
    n_scattering_stddev: int = 3
    estimate_scattering_noise: bool = True
    # pointwise_material_interactor::state
    do_energy_loss: bool = True
    do_multiple_scattering: bool = True
    # actor chain: run parameter_updater<pointwise_material_interactor>
    use_material: bool = True
    # surface_sequencer capacity
    max_surfaces: int = 64

    def to_c(self) -> capi.PropConfig:
        c = capi.PropConfig()
        c.min_stepsize = self.min_stepsize
        c.rk_error_tol = self.rk_error_tol
        c.step_constraint = self.step_constraint
        c.path_limit = self.path_limit
        c.max_rk_updates = int(self.max_rk_updates)
        c.use_mean_loss = int(self.use_mean_loss)
        c.use_eloss_gradient = int(self.use_eloss_gradient)
        c.use_field_gradient = int(self.use_field_gradient)
        c.do_covariance_transport = int(self.do_covariance_transport)
        c.min_mask_tolerance = self.min_mask_tolerance
        c.max_mask_tolerance = self.max_mask_tolerance
        c.mask_tolerance_scalor = self.mask_tolerance_scalor
        c.path_tolerance = self.path_tolerance
        c.overstep_tolerance = self.overstep_tolerance
        c.search_window[0] = int(self.search_window[0])
        c.search_window[1] = int(self.search_window[1])
        c.accumulated_error = self.accumulated_error
        c.n_scattering_stddev = int(self.n_scattering_stddev)
        c.estimate_scattering_noise = int(self.estimate_scattering_noise)
        c.do_energy_loss = int(self.do_energy_loss)
        c.do_multiple_scattering = int(self.do_multiple_scattering)
        c.use_material = int(self.use_material)
        c.max_surfaces = int(self.max_surfaces)
        return c


class Detector:
    """Flattened detector tables (host) plus a lazily created device copy."""

    def __init__(self, desc: capi.DetectorDesc, name: str, keep=None):
        self.desc = desc
        self.name = name
        self._device = None
        # tables owned by Python (from_tables) instead of the library
        self._keep = keep

    # -- builders ----------------------------------------------------------
    @classmethod
    def toy(cls, n_brl_layers: int = 4, n_edc_layers: int = 3, use_material: "bool | str" = True) -> "Detector":
        """detray's toy detector (build_toy_detector.hpp): ``use_material``
        True for silicon slabs on the modules and a beryllium beampipe,
        ``"maps"`` for material maps on every portal and the beampipe
        (toy_det_config::use_material_maps), False for none."""
        lib = capi.load()
        d = capi.DetectorDesc()
        if use_material not in (True, False, "maps"):
            raise ValueError("use_material is True, False or 'maps'")
        cfg = capi.ToyConfig(n_brl_layers, n_edc_layers, 2 if use_material == "maps" else int(use_material))
        capi.check(lib.dtr_build_toy_detector(C.byref(cfg), C.byref(d)), "dtr_build_toy_detector")
        return cls(d, f"toy_{n_brl_layers}brl_{n_edc_layers}edc")

    @classmethod
    def wire_chamber(cls, n_layers: int = 10, half_z: float = 1000.0, first_layer_inner_radius: float = 500.0,
                     cell_size: float = 10.0, stereo_angle: float = 0.05, mat_radius: float = 0.015,
                     square_cells: bool = True) -> "Detector":
        """detray's wire chamber (build_wire_chamber.hpp): drift cells
        (line_square) or straw tubes with tungsten rods, one wire layer per
        cylinder volume, 100 x 1 phi/z grids of up to 3 wires per bin."""
        lib = capi.load()
        cfg = capi.WireChamberConfig(n_layers, int(square_cells), first_layer_inner_radius, half_z, cell_size,
                                     stereo_angle, mat_radius)
        d = capi.DetectorDesc()
        capi.check(lib.dtr_build_wire_chamber(C.byref(cfg), C.byref(d)), "dtr_build_wire_chamber")
        return cls(d, "wire_chamber")

    @classmethod
    def telescope(cls, positions: Sequence[float], half_x: float = 500.0, half_y: float = 500.0,
                  thickness: float = 0.15, envelope: float = 60.0,
                  volume_material: "str | capi.MaterialSlab | None" = None) -> "Detector":
        """Telescope of rectangles along the z-axis pilot ray
        (build_telescope_detector.hpp; tel_det_config::envelope sets the
        portal box margin; tel_det_config::volume_material, by name of
        predefined_materials.hpp or as a slab record, fills the volume)."""
        lib = capi.load()
        pos = (C.c_double * len(positions))(*positions)
        vm = None
        if volume_material is not None:
            vm = material(volume_material) if isinstance(volume_material, str) else volume_material
        cfg = capi.TelescopeConfig(len(positions), 0, pos, half_x, half_y, thickness, envelope,
                                   C.pointer(vm) if vm is not None else None)
        d = capi.DetectorDesc()
        capi.check(lib.dtr_build_telescope_detector(C.byref(cfg), C.byref(d)),
                   "dtr_build_telescope_detector")
        return cls(d, f"telescope_{len(positions)}")

    @classmethod
    def from_tables(cls, volumes, surfaces, transforms, masks, materials=(), grids=(), bf_surfaces=(),
                    grid_bins=(), name: str = "tables", material_maps=()) -> "Detector":
        """Detector from explicit tables (sequences of the capi structures
        and of surface indices), e.g. for geometries no emitter builds; a
        surface links a map with material = capi.MATERIAL_MAP | index."""
        def arr(ctype, items):
            items = list(items)
            return (ctype * max(len(items), 1))(*items), len(items)
        keep = []
        d = capi.DetectorDesc()
        for field, ctype, items in (("volumes", capi.Volume, volumes), ("surfaces", capi.Surface, surfaces),
                                    ("transforms", capi.Transform, transforms), ("masks", capi.Mask, masks),
                                    ("materials", capi.MaterialSlab, materials), ("grids", capi.Grid, grids),
                                    ("bf_surfaces", C.c_uint32, bf_surfaces), ("grid_bins", C.c_uint32, grid_bins),
                                    ("material_maps", capi.MaterialMap, material_maps)):
            a, n = arr(ctype, items)
            keep.append(a)
            setattr(d, field, C.cast(a, C.POINTER(ctype)))
            count = {"volumes": "n_volumes", "surfaces": "n_surfaces", "transforms": "n_transforms",
                     "masks": "n_masks", "materials": "n_materials", "grids": "n_grids", "bf_surfaces": "n_bf",
                     "grid_bins": "n_bins", "material_maps": "n_material_maps"}[field]
            setattr(d, count, n)
        return cls(d, name, keep=keep)

    @classmethod
    def from_json(cls, geometry: "os.PathLike | str", grids: "os.PathLike | str | None" = None,
                  material: "os.PathLike | str | None" = None, name: Optional[str] = None,
                  material_maps: "os.PathLike | str | None" = None) -> "Detector":
        """Read detray's JSON detector files (``<name>_geometry.json`` and,
        optionally, ``<name>_surface_grids.json``,
        ``<name>_homogeneous_material.json`` and
        ``<name>_material_maps.json``) as the reference's io::read_detector
        builds them (dtr_read_detector_json)."""
        lib = capi.load()
        d = capi.DetectorDesc()
        msg = C.create_string_buffer(512)
        enc = lambda p: None if p is None else os.fsencode(p)
        code = lib.dtr_read_detector_json(enc(geometry), enc(grids), enc(material), enc(material_maps), C.byref(d),
                                          msg, len(msg))
        if code != capi.DTR_SUCCESS:
            raise capi.DetrayError(f"dtr_read_detector_json failed: {capi.STATUS_NAMES.get(code, code)}: "
                                   f"{msg.value.decode(errors='replace')}")
        if name is None:
            stem = Path(os.fsdecode(geometry)).name
            name = stem[:-len("_geometry.json")] if stem.endswith("_geometry.json") else Path(stem).stem
        return cls(d, name)

    def write_json(self, directory: "os.PathLike | str" = ".", name: Optional[str] = None,
                   grids: bool = True, material: bool = True,
                   volume_names: Optional[Sequence[str]] = None) -> dict:
        """Write the tables as detray JSON files into ``directory``
        (io::write_detector file names, the material-map file when a surface
        carries a map); returns {"geometry": path, ...}."""
        lib = capi.load()
        name = name or self.name
        out = Path(directory)
        files = {"geometry": out / f"{name}_geometry.json"}
        if grids:
            files["surface_grids"] = out / f"{name}_surface_grids.json"
        if material:
            files["homogeneous_material"] = out / f"{name}_homogeneous_material.json"
            if self.desc.n_material_maps:
                files["material_maps"] = out / f"{name}_material_maps.json"
        names = None
        if volume_names is not None:
            if len(volume_names) != self.desc.n_volumes:
                raise ValueError("one name per volume")
            names = (C.c_char_p * len(volume_names))(*[n.encode() for n in volume_names])
        msg = C.create_string_buffer(512)
        enc = lambda k: os.fsencode(files[k]) if k in files else None
        code = lib.dtr_write_detector_json(C.byref(self.desc), name.encode(), names, enc("geometry"),
                                           enc("surface_grids"), enc("homogeneous_material"), enc("material_maps"),
                                           msg, len(msg))
        if code != capi.DTR_SUCCESS:
            raise capi.DetrayError(f"dtr_write_detector_json failed: {capi.STATUS_NAMES.get(code, code)}: "
                                   f"{msg.value.decode(errors='replace')}")
        return {k: str(v) for k, v in files.items()}

    def check(self) -> None:
        """The table checks dtr_detector_create applies, on the host
        (dtr_check_detector); raises DetrayError naming the first problem."""
        msg = C.create_string_buffer(512)
        if capi.load().dtr_check_detector(C.byref(self.desc), msg, len(msg)) != capi.DTR_SUCCESS:
            raise capi.DetrayError(f"{self.name}: {msg.value.decode(errors='replace')}")

    # -- table views -------------------------------------------------------
    def table(self, name: str):
        n = {
            "volumes": self.desc.n_volumes, "surfaces": self.desc.n_surfaces,
            "transforms": self.desc.n_transforms, "masks": self.desc.n_masks,
            "materials": self.desc.n_materials, "grids": self.desc.n_grids,
            "bf_surfaces": self.desc.n_bf, "grid_bins": self.desc.n_bins,
            "material_maps": self.desc.n_material_maps,
        }[name]
        ptr = getattr(self.desc, name)
        return [ptr[i] for i in range(n)]

    @property
    def n_surfaces(self) -> int:
        return int(self.desc.n_surfaces)

    def device(self):
        if self._device is None:
            lib = capi.load()
            h = C.c_void_p()
            capi.check(lib.dtr_detector_create(C.byref(self.desc), C.byref(h)), "dtr_detector_create")
            self._device = h
        return self._device

    def close(self) -> None:
        lib = capi.load()
        if self._device is not None:
            lib.dtr_detector_destroy(self._device)
            self._device = None
        if self.desc.n_surfaces and self._keep is None:
            lib.dtr_host_detector_free(C.byref(self.desc))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MagneticField:
    """Constant field or a regular xyz grid with trilinear interpolation."""

    def __init__(self, desc: capi.FieldDesc, values: Optional[np.ndarray] = None, name: str = ""):
        self.desc = desc
        self.values = values  # keep the host samples alive
        self.name = name
        self._device = None

    @classmethod
    def constant(cls, bx: float, by: float, bz: float) -> "MagneticField":
        d = capi.FieldDesc()
        d.kind = 0
        d.b[0], d.b[1], d.b[2] = bx, by, bz
        return cls(d, None, f"const({bx:.3g},{by:.3g},{bz:.3g})")

    @classmethod
    def solenoid(cls, b_center: float = 2.0 * TESLA, radius: float = 1200.0,
                 half_length: float = 3000.0, r_max: float = 1500.0, z_max: float = 3500.0,
                 pitch: float = 50.0) -> "MagneticField":
        """Analytic finite solenoid sampled on x, y in [-r_max, r_max],
        z in [-z_max, z_max] with the given pitch, stored as fp32 xyz."""
        lib = capi.load()
        nxy = int(round(2 * r_max / pitch)) + 1
        nz = int(round(2 * z_max / pitch)) + 1
        n = (C.c_uint32 * 3)(nxy, nxy, nz)
        mn = (C.c_double * 3)(-r_max, -r_max, -z_max)
        pt = (C.c_double * 3)(pitch, pitch, pitch)
        values = np.zeros(nxy * nxy * nz * 3, dtype=np.float32)
        capi.check(lib.dtr_sample_solenoid(b_center, radius, half_length, n, mn, pt,
                                           values.ctypes.data_as(C.POINTER(C.c_float))),
                   "dtr_sample_solenoid")
        d = capi.FieldDesc()
        d.kind = 1
        for i in range(3):
            d.n[i] = n[i]
            d.min[i] = mn[i]
            d.pitch[i] = pt[i]
        d.values = values.ctypes.data_as(C.POINTER(C.c_float))
        return cls(d, values, f"solenoid({nxy}x{nxy}x{nz})")

    @classmethod
    def grid(cls, values: np.ndarray, minimum: Sequence[float], pitch: Sequence[float],
             name: str = "grid") -> "MagneticField":
        """A sampled field: ``values`` [nx, ny, nz, 3] (converted to fp32),
        sample (0, 0, 0) at ``minimum``, spacing ``pitch`` (trilinear
        interpolation, clamped at the border)."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.ndim != 4 or v.shape[3] != 3 or min(v.shape[:3]) < 2:
            raise ValueError("values must have shape [nx, ny, nz, 3] with at least 2 samples per axis")
        d = capi.FieldDesc()
        d.kind = 1
        for i in range(3):
            d.n[i] = v.shape[i]
            d.min[i] = float(minimum[i])
            d.pitch[i] = float(pitch[i])
        d.values = v.ctypes.data_as(C.POINTER(C.c_float))
        return cls(d, v, name)

    def device(self):
        if self._device is None:
            lib = capi.load()
            h = C.c_void_p()
            capi.check(lib.dtr_field_create(C.byref(self.desc), C.byref(h)), "dtr_field_create")
            self._device = h
        return self._device

    def close(self) -> None:
        if self._device is not None:
            capi.load().dtr_field_destroy(self._device)
            self._device = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def material(name: str, thickness: float = 0.0) -> capi.MaterialSlab:
    """A material of predefined_materials.hpp by name (e.g. "silicon",
    "cesium_iodide_with_ded", "tungsten", "isobutane"), as a slab of the
    given thickness (0 for a volume material)."""
    out = capi.MaterialSlab()
    capi.check(capi.load().dtr_predefined_material(name.encode(), float(thickness), C.byref(out)),
               f"dtr_predefined_material({name!r})")
    return out


def generate_tracks(kind: str, n: int, seed: int, **kw) -> np.ndarray:
    """Seeded track generation, returns free parameters [8, n] (float64).

    kind="beam": origin, q = +-1, |p| ~ U[p_min, p_max], polar angle
    |N(0, sigma_theta)|, phi ~ U[-pi, pi).
    kind="toy": eta ~ U[eta_min, eta_max], phi ~ U, pT log-uniform
    [pt_min, pt_max], vertex x,y ~ N(0, sigma_xy), z ~ N(0, sigma_z).
    """
    lib = capi.load()
    g = capi.TrackGenConfig()
    g.seed = seed
    g.n = n
    if kind == "beam":
        g.kind = 0
        g.p_min = kw.get("p_min", 1.0)
        g.p_max = kw.get("p_max", 10.0)
        g.sigma_theta = kw.get("sigma_theta", 0.1)
    elif kind == "toy":
        g.kind = 1
        g.eta_min = kw.get("eta_min", -2.5)
        g.eta_max = kw.get("eta_max", 2.5)
        g.pt_min = kw.get("pt_min", 1.0)
        g.pt_max = kw.get("pt_max", 100.0)
        g.sigma_xy = kw.get("sigma_xy", 0.01)
        g.sigma_z = kw.get("sigma_z", 50.0)
    else:
        raise ValueError(kind)
    out = np.zeros((8, n), dtype=np.float64)
    capi.check(lib.dtr_generate_tracks(C.byref(g), out.ctypes.data_as(C.POINTER(C.c_double))),
               "dtr_generate_tracks")
    return out


@dataclass
class LaunchConfig:
    """Arithmetic and schedule of a batched launch (``dtr_launch_config``).

    precision: "fp64" or "mixed" (FP32 Runge-Kutta stages, field, transport
    and material noise terms; FP64 track state, path length, step-size error
    estimate, mean energy loss and navigation).  schedule: "grid" (one work-item per track) or
    "persistent" (resident waves over a device work queue; a wave takes new
    tracks once ``min_idle`` of its lanes are idle).  ``order`` is an optional
    int32 device tensor [n], a permutation giving the order in which tracks
    are taken (results stay in track order).  navigation: "wave" (the lanes
    that need a volume search pool their candidates and the wave evaluates
    them together), "lane" (each lane on its own, grid schedule only) or
    "gather" (each lane on its own, and a lane that needs a volume search
    waits until every running lane of its wave does, so the wave searches
    together; grid schedule only).  ``xcd_chunk``: consecutive blocks of the
    order run on one XCD at a time (grid schedule, lane or gather
    navigation; 0 = hardware order)."""

    precision: str = "mixed"
    schedule: str = "persistent"
    block_size: int = 256
    waves_per_simd: int = 2
    tracks_per_claim: int = 64
    order: "object" = None
    min_idle: int = 1
    navigation: str = "wave"
    xcd_chunk: int = 64

    def to_c(self, queue_ptr: Optional[int]) -> capi.LaunchConfig:
        c = capi.LaunchConfig()
        c.precision = {"fp64": capi.PRECISION_FP64, "mixed": capi.PRECISION_MIXED}[self.precision]
        c.schedule = {"grid": capi.SCHEDULE_GRID, "persistent": capi.SCHEDULE_PERSISTENT}[self.schedule]
        c.block_size = int(self.block_size)
        c.waves_per_simd = int(self.waves_per_simd)
        c.tracks_per_claim = int(self.tracks_per_claim)
        c.min_idle = int(self.min_idle)
        c.navigation = {"wave": capi.NAVIGATION_WAVE, "lane": capi.NAVIGATION_LANE,
                        "gather": capi.NAVIGATION_GATHER}[self.navigation]
        c.xcd_chunk = int(self.xcd_chunk)
        c.order = self.order.data_ptr() if self.order is not None else None
        c.queue = queue_ptr
        return c


@dataclass
class PropagationResult:
    """Per-track outputs (device tensors, SoA)."""

    bound_params: "object"   # [6, n] float64
    bound_cov: "object"      # [21, n] float64 (upper triangle)
    free_params: "object"    # [8, n]
    path_length: "object"    # [n]
    surfaces: "object"       # [max_surfaces, n] uint32 (as int32 storage)
    n_surfaces: "object"     # [n]
    n_steps: "object"        # [n]
    n_trials: "object"       # [n]
    status: "object"         # [n] uint8
    trace: "object" = None   # StepTrace of the step_tracer, if it ran


@dataclass
class StepTrace:
    """step_tracer output (step_tracer.hpp: step_data per parameter update):
    entry k of track i at [..., k, i]; ``n`` [n] entries per track (may
    exceed the capacity: later entries are counted, not stored)."""

    n: "object"              # [n] int32
    surface: "object"        # [cap, n] int32
    step_size: "object"      # [cap, n] float64
    path_length: "object"    # [cap, n] float64
    n_trials: "object"       # [cap, n] int32
    free_params: "object"    # [8, cap, n] float64, the stepping state on the surface
    bound_params: "object"   # [6, cap, n] float64, after the material interactor
    jacobian: "object" = None  # [64, cap, n] float64, free transport Jacobian (row-major 8x8)


@dataclass
class Actors:
    """The optional actors of the chain and their states
    (``dtr_actor_states``).  The chain is pathlimit_aborter,
    parameter_updater<pointwise_material_interactor, step_tracer?>,
    surface_sequencer, momentum_aborter?, target_aborter?.

    momentum_aborter: stop a track once its pT <= ``min_pT`` or p <=
    ``min_p`` (aborters.hpp:80-143).  ``target_surface``: per-track target
    surface indices [n] (device int32 for the GPU, host for the oracle);
    the target_aborter stops a track on reaching it (aborters.hpp:146-175).
    ``step_tracer``: entries per track of the step tracer (0: off),
    ``trace_jacobian`` adds the transport Jacobian to them."""

    momentum_aborter: bool = False
    min_p: float = 10.0 * MEV
    min_pT: float = 10.0 * MEV
    target_surface: "object" = None
    step_tracer: int = 0
    trace_jacobian: bool = False

    def bits(self) -> int:
        return ((capi.ACTOR_MOMENTUM_ABORTER if self.momentum_aborter else 0)
                | (capi.ACTOR_TARGET_ABORTER if self.target_surface is not None else 0)
                | (capi.ACTOR_STEP_TRACER if self.step_tracer > 0 else 0))

    def allocate_trace(self, n: int, zeros):
        """StepTrace arrays for n tracks made by ``zeros(shape, kind)`` (kind
        "f8" or "i4"), or None without the step tracer."""
        if self.step_tracer <= 0:
            return None
        m = int(self.step_tracer)
        return StepTrace(zeros((n,), "i4"), zeros((m, n), "i4"), zeros((m, n), "f8"), zeros((m, n), "f8"),
                         zeros((m, n), "i4"), zeros((8, m, n), "f8"), zeros((6, m, n), "f8"),
                         zeros((64, m, n), "f8") if self.trace_jacobian else None)

    def to_c(self, trace: "Optional[StepTrace]", ptr) -> capi.ActorStates:
        """C states over the arrays of ``trace`` (``ptr`` gives an array's
        address)."""
        a = capi.ActorStates()
        capi.load().dtr_default_actors(C.byref(a))
        a.actors = self.bits()
        a.min_p = float(self.min_p)
        a.min_pT = float(self.min_pT)
        a.target_surface = ptr(self.target_surface) if self.target_surface is not None else None
        if trace is not None:
            a.trace = capi.StepTrace(int(self.step_tracer), 0, ptr(trace.n), ptr(trace.surface),
                                     ptr(trace.step_size), ptr(trace.path_length), ptr(trace.n_trials),
                                     ptr(trace.free_params), ptr(trace.bound_params),
                                     ptr(trace.jacobian) if trace.jacobian is not None else None)
        return a


class PropagationState:
    """The paused state of every track of a batch of ``n`` between calls of
    :meth:`Propagator.propagate_resumable` (``dtr_propagation_state``: a
    device buffer of about 0.9 KB per track, indexed by track)."""

    def __init__(self, n: int, device="cuda"):
        import torch

        self.n = int(n)
        nbytes = int(capi.load().dtr_propagation_state_size(self.n))
        self.data = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=device)
        self.started = False


class Propagator:
    """``detray::propagator`` over a batch of tracks on one GPU."""

    def __init__(self, config: Optional[PropagationConfig] = None,
                 launch: Optional[LaunchConfig] = None):
        """``launch`` None runs the plain FP64 kernel (``dtr_propagate``);
        otherwise the batched kernel (``dtr_propagate_ex``) with that launch
        configuration."""
        self.config = config or PropagationConfig()
        self.launch = launch
        self._c_config = self.config.to_c()
        self._queues = {}

    def _queue(self, device) -> int:
        import torch

        key = str(device)
        if key not in self._queues:
            self._queues[key] = torch.zeros(4, dtype=torch.int32, device=device)
        return self._queues[key].data_ptr()

    def allocate(self, n: int, device="cuda") -> PropagationResult:
        import torch

        ms = max(int(self.config.max_surfaces), 1)
        return PropagationResult(
            bound_params=torch.zeros((6, n), dtype=torch.float64, device=device),
            bound_cov=torch.zeros((21, n), dtype=torch.float64, device=device),
            free_params=torch.zeros((8, n), dtype=torch.float64, device=device),
            path_length=torch.zeros((n,), dtype=torch.float64, device=device),
            surfaces=torch.full((ms, n), -1, dtype=torch.int32, device=device),
            n_surfaces=torch.zeros((n,), dtype=torch.int32, device=device),
            n_steps=torch.zeros((n,), dtype=torch.int32, device=device),
            n_trials=torch.zeros((n,), dtype=torch.int32, device=device),
            status=torch.full((n,), 255, dtype=torch.uint8, device=device),
        )

    def _config_with(self, actors: Optional[Actors], n: int, device, out: PropagationResult):
        """The C config for one call: with optional actors, their C states
        (kept alive by the caller) and the trace arrays put into ``out``."""
        import torch

        if actors is None or actors.bits() == 0:
            return self._c_config, None
        ts = actors.target_surface
        if ts is not None and (not isinstance(ts, torch.Tensor) or not ts.is_cuda or ts.dtype != torch.int32
                               or tuple(ts.shape) != (n,)):
            raise ValueError("actors.target_surface must be an int32 CUDA tensor [n]")
        dt = {"f8": torch.float64, "i4": torch.int32}
        out.trace = actors.allocate_trace(n, lambda shape, k: torch.zeros(shape, dtype=dt[k], device=device))
        states = actors.to_c(out.trace, lambda t: t.contiguous().data_ptr())
        cfg = capi.PropConfig.from_buffer_copy(self._c_config)
        cfg.actors = C.addressof(states)
        return cfg, states

    def propagate(self, detector: Detector, field: MagneticField, tracks=None, initial_cov=None,
                  out: Optional[PropagationResult] = None, stream=None, bound_params=None,
                  start_surface=None, actors: Optional[Actors] = None) -> PropagationResult:
        """Propagate a batch of tracks.

        Free start: ``tracks`` is a device float64 tensor [8, n] of free
        parameters (the tracks start in volume 0).  Bound start:
        ``bound_params`` [6, n] float64 and ``start_surface`` [n] int32 give
        bound track parameters on a detector surface
        (propagator::state(bound_track_parameters, ...)).
        ``initial_cov`` [21, n] is the upper-triangle bound covariance.
        ``actors`` switches on optional actors (momentum and target aborters,
        step tracer; the tracer's output is ``out.trace``)."""
        import torch

        tb, rb, out, n, device, _keep = self._batches(detector, tracks, initial_cov, out, bound_params,
                                                      start_surface)
        cfg, _states = self._config_with(actors, n, device, out)
        lib = capi.load()
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        if self.launch is None:
            capi.check(lib.dtr_propagate(detector.device(), field.device(), C.byref(cfg),
                                         C.byref(tb), C.byref(rb), C.c_void_p(stream)),
                       "dtr_propagate")
            return out
        order = self.launch.order
        if order is not None and (not order.is_cuda or order.dtype != torch.int32 or order.shape != (n,)
                                  or not order.is_contiguous() or order.device != device):
            raise ValueError("launch.order must be a contiguous int32 CUDA tensor [n] on the tracks' device")
        lc = self.launch.to_c(self._queue(device) if self.launch.schedule == "persistent" else None)
        capi.check(lib.dtr_propagate_ex(detector.device(), field.device(), C.byref(cfg),
                                        C.byref(tb), C.byref(rb), C.byref(lc), C.c_void_p(stream)),
                   "dtr_propagate_ex")
        return out

    def propagate_resumable(self, detector: Detector, field: MagneticField, state: PropagationState,
                            max_steps: int, tracks=None, initial_cov=None, out: Optional[PropagationResult] = None,
                            stream=None, bound_params=None, start_surface=None,
                            actors: Optional[Actors] = None) -> PropagationResult:
        """One call of a propagation that pauses and resumes
        (``dtr_propagate_resumable``, propagator.hpp:238-250): the first call
        on ``state`` starts every track, each later call continues the tracks
        the previous one paused; a track pauses (status ``TRACK_PAUSED``)
        once it has taken ``max_steps`` steps in a call.  Pass the same
        inputs and the same ``out`` to every call; the results of a
        propagation done in several calls are those of one :meth:`propagate`
        call.  Needs a launch with the grid schedule and "lane" or "gather"
        navigation."""
        import torch

        if self.launch is None or self.launch.schedule != "grid" or self.launch.navigation == "wave":
            raise ValueError("propagate_resumable needs a grid launch with lane or gather navigation")
        tb, rb, out, n, device, _keep = self._batches(detector, tracks, initial_cov, out, bound_params,
                                                      start_surface)
        if state.n != n or state.data.device != device:
            raise ValueError(f"state holds {state.n} tracks on {state.data.device}, the batch {n} on {device}")
        if int(max_steps) < 1:
            raise ValueError("max_steps must be at least 1")
        cfg, _states = self._config_with(actors, n, device, out)
        order = self.launch.order
        if order is not None and (not order.is_cuda or order.dtype != torch.int32 or order.shape != (n,)
                                  or not order.is_contiguous() or order.device != device):
            raise ValueError("launch.order must be a contiguous int32 CUDA tensor [n] on the tracks' device")
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        lc = self.launch.to_c(None)
        st = capi.PropagationState(state.data.data_ptr(), int(max_steps), 1 if state.started else 0)
        capi.check(capi.load().dtr_propagate_resumable(detector.device(), field.device(), C.byref(cfg),
                                                       C.byref(tb), C.byref(rb), C.byref(lc), C.byref(st),
                                                       C.c_void_p(stream)), "dtr_propagate_resumable")
        state.started = True
        return out

    def propagate_direct(self, detector: Detector, field: MagneticField, sequence, n_sequence, tracks=None,
                         initial_cov=None, out: Optional[PropagationResult] = None, stream=None,
                         bound_params=None, start_surface=None,
                         actors: Optional[Actors] = None, backward: bool = False) -> PropagationResult:
        """Propagate with the direct navigator (``dtr_propagate_direct``):
        every track follows its surface sequence ``sequence`` [m, n] int32
        (``n_sequence`` [n] entries each, e.g. the ``surfaces`` /
        ``n_surfaces`` of an earlier propagation), from the first entry to
        the last, or with ``backward`` from the last to the first against
        the track direction (the reference's navigation direction
        e_backward).  The launch's precision, block size and order are used
        (grid schedule; without a launch: mixed precision, block 64)."""
        import torch

        tb, rb, out, n, device, _keep = self._batches(detector, tracks, initial_cov, out, bound_params,
                                                      start_surface)
        cfg, _states = self._config_with(actors, n, device, out)
        if (not isinstance(sequence, torch.Tensor) or sequence.dtype != torch.int32 or sequence.dim() != 2
                or sequence.shape[1] != n or not sequence.is_cuda):
            raise ValueError("sequence must be an int32 CUDA tensor [m, n]")
        if (not isinstance(n_sequence, torch.Tensor) or n_sequence.dtype != torch.int32
                or tuple(n_sequence.shape) != (n,) or not n_sequence.is_cuda):
            raise ValueError("n_sequence must be an int32 CUDA tensor [n]")
        sequence = sequence.contiguous()
        n_sequence = n_sequence.contiguous()
        launch = self.launch or LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "lane")
        if launch.order is not None and (not launch.order.is_cuda or launch.order.dtype != torch.int32
                                         or launch.order.shape != (n,) or not launch.order.is_contiguous()
                                         or launch.order.device != device):
            raise ValueError("launch.order must be a contiguous int32 CUDA tensor [n] on the tracks' device")
        seq = capi.SurfaceSequence(int(sequence.shape[0]),
                                   capi.DTR_DIRECTION_BACKWARD if backward else capi.DTR_DIRECTION_FORWARD,
                                   sequence.data_ptr(), n_sequence.data_ptr())
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        lc = launch.to_c(None)
        capi.check(capi.load().dtr_propagate_direct(detector.device(), field.device(), C.byref(cfg),
                                                    C.byref(tb), C.byref(seq), C.byref(rb), C.byref(lc),
                                                    C.c_void_p(stream)), "dtr_propagate_direct")
        return out

    def _check_out(self, out: PropagationResult, n: int, device) -> None:
        """A caller-supplied result set must hold n tracks (and
        ``max_surfaces`` sequence rows) in contiguous tensors of the right
        type on the tracks' device: the kernels write every column of it."""
        import torch

        ms = max(int(self.config.max_surfaces), 1)
        spec = {"bound_params": ((6, n), torch.float64), "bound_cov": ((21, n), torch.float64),
                "free_params": ((8, n), torch.float64), "path_length": ((n,), torch.float64),
                "n_surfaces": ((n,), torch.int32), "n_steps": ((n,), torch.int32),
                "n_trials": ((n,), torch.int32), "status": ((n,), torch.uint8)}
        for name, (shape, dtype) in spec.items():
            t = getattr(out, name)
            if (not isinstance(t, torch.Tensor) or tuple(t.shape) != shape or t.dtype != dtype
                    or t.device != device or not t.is_contiguous()):
                raise ValueError(f"out.{name} must be a contiguous {dtype} tensor {list(shape)} on {device}")
        t = out.surfaces
        if (not isinstance(t, torch.Tensor) or t.dim() != 2 or t.shape[1] != n or t.shape[0] < ms
                or t.dtype != torch.int32 or t.device != device or not t.is_contiguous()):
            raise ValueError(f"out.surfaces must be a contiguous int32 tensor [>= {ms}, {n}] on {device}")

    def _batches(self, detector, tracks, initial_cov, out, bound_params, start_surface):
        """Validated C views of the inputs and (allocated) outputs."""
        import torch

        if bound_params is not None:
            if start_surface is None:
                raise ValueError("bound_params needs start_surface")
            if (not bound_params.is_cuda or bound_params.dtype != torch.float64
                    or bound_params.dim() != 2 or bound_params.shape[0] != 6):
                raise ValueError("bound_params must be a float64 CUDA tensor [6, n]")
            bound_params = bound_params.contiguous()
            n = int(bound_params.shape[1])
            if (not start_surface.is_cuda or start_surface.dtype != torch.int32
                    or start_surface.shape != (n,)):
                raise ValueError("start_surface must be an int32 CUDA tensor [n]")
            start_surface = start_surface.contiguous()
            if int(start_surface.max().item() if n else 0) >= detector.n_surfaces or (
                    n and int(start_surface.min().item()) < 0):
                raise ValueError("start_surface out of range")
            device = bound_params.device
        else:
            if not (isinstance(tracks, torch.Tensor) and tracks.is_cuda):
                raise TypeError("tracks must be a CUDA (HIP) tensor [8, n] of float64")
            if tracks.dtype != torch.float64 or tracks.dim() != 2 or tracks.shape[0] != 8:
                raise ValueError("tracks must be float64 with shape [8, n]")
            tracks = tracks.contiguous()
            n = int(tracks.shape[1])
            device = tracks.device
        if initial_cov is not None:
            if initial_cov.shape != (21, n) or initial_cov.dtype != torch.float64 or not initial_cov.is_cuda:
                raise ValueError("initial_cov must be a float64 CUDA tensor [21, n]")
            initial_cov = initial_cov.contiguous()
        if out is None:
            out = self.allocate(n, device)
        else:
            self._check_out(out, n, device)
        tb = capi.TrackBatch(n, 0, tracks.data_ptr() if tracks is not None and bound_params is None else None,
                             initial_cov.data_ptr() if initial_cov is not None else None,
                             bound_params.data_ptr() if bound_params is not None else None,
                             start_surface.data_ptr() if start_surface is not None else None)
        rb = capi.ResultBatch(out.bound_params.data_ptr(), out.bound_cov.data_ptr(),
                              out.free_params.data_ptr(), out.path_length.data_ptr(),
                              out.surfaces.data_ptr(), out.n_surfaces.data_ptr(),
                              out.n_steps.data_ptr(), out.n_trials.data_ptr(),
                              out.status.data_ptr())
        # the tensors behind the C views stay referenced until the launch
        return tb, rb, out, n, device, (tracks, initial_cov, bound_params, start_surface)
