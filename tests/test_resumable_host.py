"""Pause and resume on the CPU: the engine's lane loop with a step budget per
call (propagate_resumable, detray_amd/csrc/device/engine.hpp, compiled for
the host by tests/native/engine_host.cpp) against the same loop in one call.

The reference's propagator pauses a track when an actor stops its heartbeat
while the navigation is alive and continues it in a later propagate() call
without initialising the navigation (propagator.hpp:238-250, 274-282); the
contract is that a propagation done in several calls gives the results of
one call bit for bit, whatever the budget, on every geometry (toy, wire
chamber, volume material), in both precisions and with noise estimation
(which reads the covariance between calls)."""

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import engine_host

KEYS = ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "n_steps",
        "n_trials", "status")


def assert_same(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("budget", [1, 3, 8])
def test_toy_in_several_calls_equals_one_call(toy7, solenoid, mixed, budget):
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 300, 43)
    ref = engine_host.propagate(toy7, solenoid, cfg, fp, mixed=mixed)
    got, calls, first = engine_host.propagate_resumable(toy7, solenoid, cfg, fp, budget, mixed=mixed)
    assert first == (ref["n_steps"] > budget).sum()
    assert calls >= int(np.ceil(ref["n_steps"].max() / budget))
    assert_same(got, ref)


@pytest.mark.parametrize("mixed", [False, True])
def test_wire_chamber_and_volume_material(solenoid, mixed):
    cfg = da.PropagationConfig(search_window=(3, 3))
    chamber = da.Detector.wire_chamber(n_layers=6, half_z=500.0)
    fp = da.generate_tracks("toy", 200, 44, eta_min=-1.0, eta_max=1.0)
    ref = engine_host.propagate(chamber, solenoid, cfg, fp, mixed=mixed)
    got, _, first = engine_host.propagate_resumable(chamber, solenoid, cfg, fp, 2, mixed=mixed)
    assert first > 0
    assert_same(got, ref)
    slab = da.material("silicon", 0.0)
    tel = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0, volume_material=slab)
    fpt = da.generate_tracks("beam", 200, 42)
    ref = engine_host.propagate(tel, da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0), cfg, fpt, mixed=mixed)
    got, _, first = engine_host.propagate_resumable(tel, da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0),
                                                    cfg, fpt, 3, mixed=mixed)
    assert first > 0
    assert_same(got, ref)


def test_paused_tracks_report_their_progress(toy7, solenoid):
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 100, 45)
    ref = engine_host.propagate(toy7, solenoid, cfg, fp)
    got, calls, first = engine_host.propagate_resumable(toy7, solenoid, cfg, fp, 5, max_calls=1)
    assert calls == 1
    paused = got["status"] == capi.TRACK_PAUSED
    assert paused.sum() == first == (ref["n_steps"] > 5).sum()
    assert (got["n_steps"][paused] == 5).all()
    assert (got["n_surfaces"][paused] <= ref["n_surfaces"][paused]).all()
    done = ~paused
    for k in KEYS:
        np.testing.assert_array_equal(got[k][..., done], ref[k][..., done], err_msg=k)
