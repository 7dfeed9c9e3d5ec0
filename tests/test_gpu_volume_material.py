"""Continuous energy loss in volume material on the device: a telescope of
ten silicon planes in a CsI-filled volume (density-effect data), 1-10 GeV
muons in a field across the beam, with the mean loss, use_eloss_gradient
and neither, against the oracle (whose volume-material stepper is pinned in
test_oracle_volume_material.py): the plain FP64 kernel and the per-lane and
gathered batched kernels (FP64 contract for FP64, the mixed-precision
contract for mixed), bound starts, the direct navigator replaying the
oracle's sequences; the pooled navigation, the persistent schedule and the
field gradient return DTR_ERROR_UNSUPPORTED on such detectors."""

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host
from tests.test_engine_host import volume_material_case
from tests.test_gpu_batched import check, run_launch
from tests.test_gpu_direct import run_direct
from tests.test_gpu_parity import assert_parity, run_both

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CONFIGS = [{}, {"use_eloss_gradient": True}, {"use_mean_loss": False}]
IDS = ["default", "eloss_gradient", "no_mean_loss"]


@pytest.mark.parametrize("kw", CONFIGS, ids=IDS)
def test_plain_kernel_matches_oracle(kw):
    det, field, cfg, fp, cov = volume_material_case(n=2048, **kw)
    got, ref = run_both(det, field, cfg, fp, initial_cov=cov)
    assert np.all(ref["n_surfaces"] == 10)
    # the volume material slowed every track down
    assert np.all(np.abs(ref["free_params"][7]) > np.abs(fp[7]) * (1.0 + 1e-3))
    assert_parity(got, ref)


@pytest.mark.parametrize("kw", CONFIGS, ids=IDS)
@pytest.mark.parametrize("precision,nav", [("fp64", "lane"), ("fp64", "gather"), ("mixed", "lane"),
                                           ("mixed", "gather")])
def test_batched_kernels_match_oracle(kw, precision, nav):
    det, field, cfg, fp, cov = volume_material_case(n=2048, **kw)
    launch = da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, nav)
    got = run_launch(det, field, cfg, fp, launch, initial_cov=cov)
    ref = oracle_host.propagate(det, field, cfg, fp, initial_cov=cov)
    check(got, ref, precision)


def test_bound_start_matches_oracle():
    det, field, cfg, fp, cov = volume_material_case(n=1024, use_eloss_gradient=True)
    rec = oracle_host.propagate(det, field, cfg, fp, initial_cov=cov)
    bp = rec["bound_params"].copy()
    # start again from the last plane, backwards along the telescope
    bp[2] += np.pi
    bp[3] = np.pi - bp[3]
    bp[4] = -bp[4]
    start = np.full(bp.shape[1], 9, dtype=np.uint32)
    ref = oracle_host.propagate(det, field, cfg, None, initial_cov=cov, bound_params=bp, start_surface=start)
    assert np.mean(ref["n_surfaces"]) > 5
    for precision in ("fp64", "mixed"):
        launch = da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "lane")
        got = run_launch(det, field, cfg, None, launch, initial_cov=cov, bound=bp, start=start)
        check(got, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_direct_navigator_matches_oracle(precision):
    det, field, cfg, fp, cov = volume_material_case(n=1024)
    rec = oracle_host.propagate(det, field, cfg, fp)
    ref = oracle_host.propagate_direct(det, field, cfg, fp, rec["surfaces"], rec["n_surfaces"])
    got = run_direct(det, field, cfg, fp, rec["surfaces"], rec["n_surfaces"], precision)
    np.testing.assert_array_equal(got["status"], ref["status"])
    np.testing.assert_array_equal(got["n_surfaces"], ref["n_surfaces"])
    if precision == "fp64":
        assert_parity(got, ref, check_cov=False)
    else:
        err = np.abs(got["bound_params"] - ref["bound_params"]) / np.maximum(np.abs(ref["bound_params"]), 1.0)
        assert err.max() <= 1e-4, err.max()


@pytest.mark.parametrize("schedule,nav", [("grid", "wave"), ("persistent", "wave")])
def test_pooled_navigation_is_unsupported(schedule, nav):
    det, field, cfg, fp, cov = volume_material_case(n=64)
    launch = da.LaunchConfig("mixed", schedule, 64, 1, 64, None, 1, nav)
    with pytest.raises(capi.DetrayError, match="UNSUPPORTED"):
        run_launch(det, field, cfg, fp, launch)


def test_field_gradient_is_unsupported():
    det, field, cfg, fp, cov = volume_material_case(n=64, use_field_gradient=True)
    with pytest.raises(capi.DetrayError, match="UNSUPPORTED"):
        run_both(det, field, cfg, fp)
