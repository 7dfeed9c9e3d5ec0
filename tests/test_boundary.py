"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/detray_amd.h declares; the Python config mirror matches the C
defaults (propagation::config)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi

HEADER = Path(__file__).resolve().parent.parent / "include" / "detray_amd.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:DTR_API\s+)?(?:int|void|size_t|const char\*)\s+(dtr_\w+)\s*\(", text, re.M)))


def test_library_exports_all_header_symbols():
    lib = capi.load()
    names = header_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(capi.HEADER_SYMBOLS)


def test_struct_sizes_match_header_layout():
    assert C.sizeof(capi.Transform) == 96
    assert C.sizeof(capi.Mask) == 64
    assert C.sizeof(capi.Surface) == 32
    assert C.sizeof(capi.Volume) == 28
    assert C.sizeof(capi.MaterialSlab) == 120
    assert C.sizeof(capi.Grid) == 56
    assert C.sizeof(capi.WireChamberConfig) == 48


def test_default_config_matches_mirror():
    lib = capi.load()
    c = capi.PropConfig()
    lib.dtr_default_config(C.byref(c))
    m = da.PropagationConfig().to_c()
    for name, _ in capi.PropConfig._fields_:
        a, b = getattr(c, name), getattr(m, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    assert np.float32(c.path_tolerance) == np.float32(1e-3)
    assert np.float32(c.overstep_tolerance) == np.float32(-1.0)


def test_version():
    assert b"gfx950" in capi.load().dtr_version()


def test_invalid_arguments_are_reported():
    lib = capi.load()
    d = capi.DetectorDesc()
    assert lib.dtr_build_toy_detector(None, C.byref(d)) == 1
    bad = capi.ToyConfig(2, 3, 1)  # endcaps need all four barrel layers
    assert lib.dtr_build_toy_detector(C.byref(bad), C.byref(d)) == 1
    assert lib.dtr_propagate(None, None, None, None, None, None) == 1


def test_default_launch_matches_mirror():
    lib = capi.load()
    c = capi.LaunchConfig()
    lib.dtr_default_launch(C.byref(c))
    m = da.LaunchConfig().to_c(None)
    for name, _ in capi.LaunchConfig._fields_:
        assert getattr(c, name) == getattr(m, name), name
    assert c.precision == capi.PRECISION_MIXED and c.schedule == capi.SCHEDULE_PERSISTENT
    assert C.sizeof(capi.LaunchConfig) == 48


def test_propagate_ex_rejects_bad_launch_before_touching_the_device():
    lib = capi.load()
    cfg = capi.PropConfig()
    lib.dtr_default_config(C.byref(cfg))
    tb = capi.TrackBatch(0, 0, None, None, None, None)
    rb = capi.ResultBatch()
    # null handles are rejected first
    assert lib.dtr_propagate_ex(None, None, C.byref(cfg), C.byref(tb), C.byref(rb), None, None) == 1
    # waves_per_simd beyond the two register budgets the kernels are built
    # for is rejected (it used to alias the W = 2 kernel)
    lc = capi.LaunchConfig()
    lib.dtr_default_launch(C.byref(lc))
    lc.queue = C.c_void_p(8)
    fake = C.c_void_p(8)  # never dereferenced: an empty batch returns after the launch checks
    for w, rc in ((1, 0), (2, 0), (3, 1), (4, 1), (32, 1)):
        lc.waves_per_simd = w
        assert lib.dtr_propagate_ex(fake, fake, C.byref(cfg), C.byref(tb), C.byref(rb), C.byref(lc), None) == rc, w


def test_propagate_resumable_checks_before_touching_the_device():
    """State size (98 doubles and 25 words per track) and the argument
    checks of dtr_propagate_resumable, all before any device work."""
    lib = capi.load()
    assert lib.dtr_propagation_state_size(1000) == 1000 * (98 * 8 + 25 * 4)
    cfg = capi.PropConfig()
    lib.dtr_default_config(C.byref(cfg))
    tb = capi.TrackBatch(0, 0, None, None, None, None)
    rb = capi.ResultBatch()
    lc = capi.LaunchConfig()
    lib.dtr_default_launch(C.byref(lc))
    lc.schedule, lc.navigation = capi.SCHEDULE_GRID, capi.NAVIGATION_GATHER
    fake = C.c_void_p(8)  # never dereferenced: an empty batch returns after the checks
    st = capi.PropagationState(None, 4, 0)
    call = lambda s: lib.dtr_propagate_resumable(fake, fake, C.byref(cfg), C.byref(tb), C.byref(rb), C.byref(lc),
                                                  s, None)
    assert call(C.byref(st)) == 0
    assert call(None) == 1
    assert call(C.byref(capi.PropagationState(None, 0, 0))) == 1  # a budget of no steps
    assert call(C.byref(capi.PropagationState(None, 4, 2))) == 1
    lc.navigation = capi.NAVIGATION_WAVE
    assert call(C.byref(st)) == 5  # DTR_ERROR_UNSUPPORTED: pooled navigation
    lc.navigation, lc.schedule = capi.NAVIGATION_LANE, capi.SCHEDULE_PERSISTENT
    assert call(C.byref(st)) == 5


def _copy_tables(det):
    """Python-owned copies of a detector's tables (for corruption tests)."""
    return {name: [type(x).from_buffer_copy(x) if hasattr(x, "_fields_") else int(x) for x in det.table(name)]
            for name in ("volumes", "surfaces", "transforms", "masks", "materials", "grids", "bf_surfaces",
                         "grid_bins")}


def _create(tables):
    d = da.Detector.from_tables(**tables)
    h = C.c_void_p()
    return capi.load().dtr_detector_create(C.byref(d.desc), C.byref(h))


def _corrupt_cases():
    yield "mask link past the volumes", lambda t: setattr(t["masks"][5], "volume_link", len(t["volumes"]))
    yield "volume grid out of range", lambda t: setattr(t["volumes"][7], "grid", len(t["grids"]))
    yield "volume brute-force range", lambda t: setattr(t["volumes"][3], "bf_end", len(t["bf_surfaces"]) + 1)
    yield "volume surface range", lambda t: setattr(t["volumes"][2], "sf_end", len(t["surfaces"]) + 5)
    yield "volume transform", lambda t: setattr(t["volumes"][1], "transform", len(t["transforms"]))
    yield "surface volume", lambda t: setattr(t["surfaces"][9], "volume", len(t["volumes"]))
    yield "surface material", lambda t: setattr(t["surfaces"][9], "material", len(t["materials"]))
    yield "grid bins", lambda t: setattr(t["grids"][0], "bins", len(t["grid_bins"]))
    yield "bin surface", lambda t: t["grid_bins"].__setitem__(3, len(t["surfaces"]))
    yield "brute-force surface", lambda t: t["bf_surfaces"].__setitem__(0, len(t["surfaces"]) + 1)


@pytest.mark.parametrize("case", [c[0] for c in _corrupt_cases()])
def test_detector_create_rejects_inconsistent_tables(toy3, case):
    """Every index the kernels follow is validated on the host before any
    upload: a table that points out of range is DTR_ERROR_INVALID_ARGUMENT
    (here without a device; consistent tables reach the device check)."""
    mutate = dict(_corrupt_cases())[case]
    t = _copy_tables(toy3)
    rc_good = _create(t)
    assert rc_good in (0, 4)  # success with a GPU, DTR_ERROR_NO_DEVICE without
    mutate(t)
    assert _create(t) == 1, case


def test_caller_out_is_validated():
    """A caller-supplied result set must fit the batch: wrong shapes, types,
    too few sequence rows or non-contiguous tensors are rejected before any
    launch (checked here with host tensors)."""
    torch = pytest.importorskip("torch")
    prop = da.Propagator(da.PropagationConfig(max_surfaces=16))
    cpu = torch.device("cpu")
    good = prop.allocate(8, cpu)
    prop._check_out(good, 8, cpu)
    with pytest.raises(ValueError, match="bound_params"):
        prop._check_out(prop.allocate(4, cpu), 8, cpu)
    small = da.Propagator(da.PropagationConfig(max_surfaces=8)).allocate(8, cpu)
    with pytest.raises(ValueError, match="surfaces"):
        prop._check_out(small, 8, cpu)
    bad = prop.allocate(8, cpu)
    bad.status = bad.status.to(torch.int32)
    with pytest.raises(ValueError, match="status"):
        prop._check_out(bad, 8, cpu)
    strided = prop.allocate(8, cpu)
    strided.bound_cov = torch.zeros((8, 21), dtype=torch.float64).t()
    with pytest.raises(ValueError, match="bound_cov"):
        prop._check_out(strided, 8, cpu)
