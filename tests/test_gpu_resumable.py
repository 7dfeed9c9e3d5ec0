"""Pause and resume on the device (dtr_propagate_resumable,
propagator.hpp:238-250): a propagation done in several calls with a step
budget per call equals one dtr_propagate_ex call with the same launch bit
for bit (and so the oracle at the batch contracts), for the per-lane and
gathered kernels, both precisions, the toy in the solenoid map (sorted
track order, noise estimation on) and the wire chamber; paused tracks
report their progress; unsupported launches are rejected."""

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host
from tests.test_gpu_batched import check

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

KEYS = ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "n_steps",
        "n_trials", "status")


def host(res):
    return {k: (getattr(res, k).cpu().numpy().view(np.uint32)
                if k in ("surfaces", "n_surfaces", "n_steps", "n_trials") else getattr(res, k).cpu().numpy())
            for k in KEYS}


def resumable(det, field, cfg, tracks, launch, budget, max_calls=10000):
    prop = da.Propagator(cfg, launch)
    state = da.PropagationState(tracks.shape[1], tracks.device)
    out = prop.allocate(tracks.shape[1], tracks.device)
    calls, first = 0, None
    while calls < max_calls:
        prop.propagate_resumable(det, field, state, budget, tracks, out=out)
        paused = int((out.status == capi.TRACK_PAUSED).sum().item())
        first = paused if first is None else first
        calls += 1
        if paused == 0:
            break
    torch.cuda.synchronize()
    return host(out), calls, first


@pytest.mark.parametrize("precision,nav", [("mixed", "gather"), ("fp64", "lane"), ("mixed", "lane"),
                                           ("fp64", "gather")])
@pytest.mark.parametrize("budget", [1, 5])
def test_toy_in_several_calls_equals_one_call(toy7, solenoid, precision, nav, budget):
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 8192, 44)
    tracks = torch.from_numpy(fp).cuda()
    from detray_amd.schedule import sort_order

    launch = da.LaunchConfig(precision, "grid", 64, 2, 64, sort_order(tracks, "cells_phi"), 1, nav, 64)
    one = da.Propagator(cfg, launch).propagate(toy7, solenoid, tracks)
    torch.cuda.synchronize()
    ref = host(one)
    got, calls, first = resumable(toy7, solenoid, cfg, tracks, launch, budget)
    assert first == int((ref["n_steps"] > budget).sum())
    assert calls >= int(np.ceil(ref["n_steps"].max() / budget))
    for k in KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    if budget == 5:
        check(got, oracle_host.propagate(toy7, solenoid, cfg, fp), precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wire_chamber(solenoid, precision):
    det = da.Detector.wire_chamber(half_z=500.0)
    cfg = da.PropagationConfig(search_window=(3, 3))
    tracks = torch.from_numpy(da.generate_tracks("toy", 4000, 44, eta_min=-1.0, eta_max=1.0)).cuda()
    launch = da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "gather", 64)
    one = da.Propagator(cfg, launch).propagate(det, solenoid, tracks)
    torch.cuda.synchronize()
    ref = host(one)
    got, _, first = resumable(det, solenoid, cfg, tracks, launch, 3)
    assert first > 0
    for k in KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_paused_tracks_report_their_progress(toy7, solenoid):
    cfg = da.PropagationConfig(search_window=(3, 3))
    tracks = torch.from_numpy(da.generate_tracks("toy", 2048, 45)).cuda()
    launch = da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "gather", 64)
    ref = host(da.Propagator(cfg, launch).propagate(toy7, solenoid, tracks))
    got, calls, _ = resumable(toy7, solenoid, cfg, tracks, launch, 7, max_calls=1)
    assert calls == 1
    paused = got["status"] == capi.TRACK_PAUSED
    assert paused.sum() == (ref["n_steps"] > 7).sum() > 0
    assert (got["n_steps"][paused] == 7).all()
    assert (got["n_surfaces"][paused] <= ref["n_surfaces"][paused]).all()
    for k in KEYS:
        np.testing.assert_array_equal(got[k][..., ~paused], ref[k][..., ~paused], err_msg=k)


def test_unsupported_and_invalid_launches(toy7, solenoid):
    cfg = da.PropagationConfig(search_window=(3, 3))
    tracks = torch.from_numpy(da.generate_tracks("toy", 256, 46)).cuda()
    state = da.PropagationState(256, tracks.device)
    for launch in (da.LaunchConfig("mixed", "persistent", 64, 2, 64, None, 1, "wave"),
                   da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "wave")):
        with pytest.raises(ValueError):
            da.Propagator(cfg, launch).propagate_resumable(toy7, solenoid, state, 4, tracks)
    good = da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "gather")
    with pytest.raises(ValueError):
        da.Propagator(cfg, good).propagate_resumable(toy7, solenoid, state, 0, tracks)
    with pytest.raises(ValueError):
        da.Propagator(cfg, good).propagate_resumable(toy7, solenoid, da.PropagationState(128, tracks.device), 4,
                                                     tracks)
    grad = da.PropagationConfig(search_window=(3, 3), use_field_gradient=True)
    with pytest.raises(capi.DetrayError, match="UNSUPPORTED"):
        da.Propagator(grad, good).propagate_resumable(toy7, solenoid, state, 4, tracks)
