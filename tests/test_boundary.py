"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/detray_amd.h declares; the Python config mirror matches the C
defaults (propagation::config)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np

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
    assert C.sizeof(capi.MaterialSlab) == 80
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
