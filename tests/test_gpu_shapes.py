"""Annulus and placed-cylinder surfaces on the device: the plain FP64 kernel
against the oracle at the FP64 contract (telescopes of annuli and of
cylinders across the beam, with material, constant field with and without
noise estimation); the batched kernels (grid schedule, per-lane navigation)
at the FP64 and mixed-precision contracts, with per-lane and pooled (wave)
navigation and the persistent schedule, and the direct navigator against the
oracle's."""

import pytest

import detray_amd as da
from tests import oracle_host
from tests.test_gpu_batched import check, run_launch
from tests.test_gpu_direct import run_direct
from tests.test_gpu_parity import assert_parity, run_both
from tests.test_shapes import _modified_telescope

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("noise", [False, True])
@pytest.mark.parametrize("kind,n", [("annulus", 6), ("cylinder", 3)])
def test_plain_kernel_matches_oracle(kind, n, noise):
    det = _modified_telescope(kind, n=n)
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(estimate_scattering_noise=noise)
    got, ref = run_both(det, field, cfg, da.generate_tracks("beam", 4000, 42, sigma_theta=0.1))
    assert ref["n_surfaces"].mean() > 3
    assert_parity(got, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind,n", [("annulus", 6), ("cylinder", 3)])
def test_batched_kernel_matches_oracle(kind, n, precision):
    det = _modified_telescope(kind, n=n)
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(estimate_scattering_noise=True)
    fp = da.generate_tracks("beam", 4000, 42, sigma_theta=0.1)
    ref = oracle_host.propagate(det, field, cfg, fp, threads=oracle_host.default_threads())
    got = run_launch(det, field, cfg, fp, da.LaunchConfig(precision, "grid", 256, 2, 64, None, 1, "lane"))
    check(got, ref, precision)
    got = run_launch(det, field, cfg, fp, da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "gather"))
    assert ref["n_surfaces"].mean() > 3
    check(got, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind,n", [("annulus", 6), ("cylinder", 3)])
def test_direct_navigator_matches_oracle(kind, n, precision):
    det = _modified_telescope(kind, n=n)
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(estimate_scattering_noise=False)
    fp = da.generate_tracks("beam", 4000, 43, sigma_theta=0.1)
    rec = oracle_host.propagate(det, field, cfg, fp, threads=oracle_host.default_threads())
    ref = oracle_host.propagate_direct(det, field, cfg, fp, rec["surfaces"], rec["n_surfaces"],
                                       threads=oracle_host.default_threads())
    got = run_direct(det, field, cfg, fp, rec["surfaces"], rec["n_surfaces"], precision)
    check(got, ref, precision)


@pytest.mark.parametrize("kind,n", [("annulus", 6), ("cylinder", 3)])
def test_pooled_navigation_and_persistent_schedule(kind, n):
    det = _modified_telescope(kind, n=n)
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(estimate_scattering_noise=True)
    fp = da.generate_tracks("beam", 4000, 44, sigma_theta=0.1)
    ref = oracle_host.propagate(det, field, cfg, fp, threads=oracle_host.default_threads())
    for precision in ("fp64", "mixed"):
        for launch in (da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "wave"),
                       da.LaunchConfig(precision, "persistent", 256, 2, 64, None, 1, "wave")):
            check(run_launch(det, field, cfg, fp, launch), ref, precision)


def test_cylinders_in_a_grid_bin():
    """Placed cylinders in a multi-surface grid bin: the pooled navigation
    hands them to their owner lane, which inserts both crossings in order."""
    from tests.test_shapes import _gridded_cylinder_telescope

    det = _gridded_cylinder_telescope(3)
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(0, 0), estimate_scattering_noise=True)
    fp = da.generate_tracks("beam", 4000, 45, sigma_theta=0.1)
    ref = oracle_host.propagate(det, field, cfg, fp, threads=oracle_host.default_threads())
    assert ref["n_surfaces"].mean() > 3
    for precision in ("fp64", "mixed"):
        for launch in (da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "lane"),
                       da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "wave"),
                       da.LaunchConfig(precision, "persistent", 256, 2, 64, None, 1, "wave")):
            check(run_launch(det, field, cfg, fp, launch), ref, precision)
