"""use_field_gradient on the device: the plain FP64 kernel and the batched
grid/per-lane kernels (FP64, mixed) against the oracle with the same flag,
in the solenoid map (toy detector) and in a sampled field with a linear
gradient (telescope).  Same tolerances as the kernels without the flag."""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host
from tests.test_field_gradient import gradient_field
from tests.test_gpu_batched import check, run_launch
from tests.test_gpu_parity import _cov_full, assert_parity, run_both

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

TOY = dict(search_window=(3, 3), estimate_scattering_noise=False)


def test_plain_kernel_solenoid(toy7, solenoid):
    cfg = da.PropagationConfig(use_field_gradient=True, **TOY)
    fp = da.generate_tracks("toy", 2000, 44)
    got, ref = run_both(toy7, solenoid, cfg, fp)
    assert_parity(got, ref)
    # the flag reaches the device: the covariance differs from the one without it
    plain = oracle_host.propagate(toy7, solenoid, da.PropagationConfig(**TOY), fp)
    d = np.abs(_cov_full(got["bound_cov"]) - _cov_full(plain["bound_cov"])).max()
    assert d > 0.0


@pytest.mark.parametrize("kind", ["Bx(x)", "Bz(y)"])
def test_plain_kernel_gradient_field(kind):
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0)
    cfg = da.PropagationConfig(step_constraint=50.0, use_field_gradient=True)
    fp = da.generate_tracks("beam", 1000, 42)
    got, ref = run_both(det, gradient_field(kind), cfg, fp)
    assert_parity(got, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_kernels(toy7, solenoid, precision):
    cfg = da.PropagationConfig(use_field_gradient=True, **TOY)
    fp = da.generate_tracks("toy", 2000, 45)
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    for block, wps, nav in ((64, 2, "lane"), (256, 1, "lane"), (64, 2, "gather")):
        launch = da.LaunchConfig(precision, "grid", block, wps, 64, None, 1, nav)
        check(run_launch(toy7, solenoid, cfg, fp, launch), ref, precision)


def test_batched_kernel_gradient_field_with_noise():
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0)
    field = gradient_field("Bz(y)")
    cfg = da.PropagationConfig(step_constraint=50.0, use_field_gradient=True)
    fp = da.generate_tracks("beam", 1000, 42)
    ref = oracle_host.propagate(det, field, cfg, fp)
    for precision in ("fp64", "mixed"):
        launch = da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "lane")
        check(run_launch(det, field, cfg, fp, launch), ref, precision)


@pytest.mark.parametrize("schedule,nav", [("persistent", "wave"), ("grid", "wave"), ("persistent", "lane")])
def test_other_launches_are_rejected(toy7, solenoid, schedule, nav):
    cfg = da.PropagationConfig(use_field_gradient=True, **TOY)
    fp = torch.from_numpy(da.generate_tracks("toy", 64, 45)).cuda()
    launch = da.LaunchConfig("mixed", schedule, 64, 1, 64, None, 1, nav)
    with pytest.raises(RuntimeError):
        da.Propagator(cfg, launch).propagate(toy7, solenoid, fp)


def test_eloss_gradient_flag(toy7, solenoid):
    """No volume material: use_eloss_gradient changes nothing."""
    fp = torch.from_numpy(da.generate_tracks("toy", 512, 43)).cuda()
    a = da.Propagator(da.PropagationConfig(**TOY)).propagate(toy7, solenoid, fp)
    b = da.Propagator(da.PropagationConfig(use_eloss_gradient=True, **TOY)).propagate(toy7, solenoid, fp)
    torch.cuda.synchronize()
    assert torch.equal(a.bound_params, b.bound_params)
    assert torch.equal(a.bound_cov, b.bound_cov)
