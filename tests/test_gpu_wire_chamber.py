"""Wire chamber on the device (drift cells and straw tubes, tungsten rods,
grid bins of up to 3 wires): the plain FP64 kernel and the batched kernels
(grid schedule, per-lane navigation) against the oracle, at the FP64
contract for FP64 and at the mixed-precision contract of the batched tests
for mixed precision, also with the pooled (wave) navigation and the
persistent schedule; the direct navigator replaying the oracle's sequences
against the oracle's direct navigator; use_field_gradient (line Jacobians
through the position columns) on the plain path and the per-lane batched
kernels, DTR_ERROR_UNSUPPORTED with the pooled navigation."""

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host
from tests.test_gpu_batched import check, run_launch
from tests.test_gpu_direct import run_direct
from tests.test_gpu_parity import assert_parity, run_both

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chamber():
    return da.Detector.wire_chamber(half_z=500.0)


@pytest.mark.parametrize("noise", [False, True])
@pytest.mark.parametrize("field_kind", ["const", "solenoid"])
def test_plain_kernel_matches_oracle(chamber, solenoid, noise, field_kind):
    field = da.MagneticField.constant(0, 0, 2 * da.TESLA) if field_kind == "const" else solenoid
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=noise)
    fp = da.generate_tracks("toy", 4000, 44, eta_min=-1.0, eta_max=1.0)
    got, ref = run_both(chamber, field, cfg, fp)
    assert ref["n_surfaces"].mean() > 5
    assert_parity(got, ref)


def test_straw_tubes(solenoid):
    det = da.Detector.wire_chamber(n_layers=4, half_z=500.0, square_cells=False)
    cfg = da.PropagationConfig(search_window=(3, 3))
    got, ref = run_both(det, solenoid, cfg, da.generate_tracks("toy", 2000, 45, eta_min=-1.0, eta_max=1.0))
    assert_parity(got, ref)


def test_json_read_chamber_propagates_identically(tmp_path, chamber, solenoid):
    f = chamber.write_json(tmp_path, "wire_chamber")
    det = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    cfg = da.PropagationConfig(search_window=(3, 3))
    tracks = torch.from_numpy(da.generate_tracks("toy", 2000, 46, eta_min=-1.0, eta_max=1.0)).cuda()
    a = da.Propagator(cfg).propagate(chamber, solenoid, tracks)
    b = da.Propagator(cfg).propagate(det, solenoid, tracks)
    torch.cuda.synchronize()
    for k in ("bound_params", "bound_cov", "surfaces", "n_surfaces", "status"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("field_kind", ["const", "solenoid"])
def test_batched_kernel_matches_oracle(chamber, solenoid, precision, field_kind):
    field = da.MagneticField.constant(0, 0, 2 * da.TESLA) if field_kind == "const" else solenoid
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=True)
    fp = da.generate_tracks("toy", 4000, 48, eta_min=-1.0, eta_max=1.0)
    ref = oracle_host.propagate(chamber, field, cfg, fp, threads=oracle_host.default_threads())
    order = torch.from_numpy(np.random.default_rng(3).permutation(fp.shape[1]).astype(np.int32)).cuda()
    for launch in (da.LaunchConfig(precision, "grid", 256, 2, 64, None, 1, "lane"),
                   da.LaunchConfig(precision, "grid", 64, 1, 64, order, 1, "lane"),
                   da.LaunchConfig(precision, "grid", 64, 2, 64, order, 1, "gather")):
        got = run_launch(chamber, field, cfg, fp, launch)
        assert ref["n_surfaces"].mean() > 5
        check(got, ref, precision)


def test_batched_straw_tubes(solenoid):
    det = da.Detector.wire_chamber(n_layers=4, half_z=500.0, square_cells=False)
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 2000, 45, eta_min=-1.0, eta_max=1.0)
    ref = oracle_host.propagate(det, solenoid, cfg, fp, threads=oracle_host.default_threads())
    for precision in ("fp64", "mixed"):
        check(run_launch(det, solenoid, cfg, fp, da.LaunchConfig(precision, "grid", 256, 2, 64, None, 1, "lane")),
              ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_direct_navigator_matches_oracle(chamber, solenoid, precision):
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = da.generate_tracks("toy", 4000, 49, eta_min=-1.0, eta_max=1.0)
    rec = oracle_host.propagate(chamber, solenoid, cfg, fp, threads=oracle_host.default_threads())
    ref = oracle_host.propagate_direct(chamber, solenoid, cfg, fp, rec["surfaces"], rec["n_surfaces"],
                                       threads=oracle_host.default_threads())
    assert np.array_equal(ref["surfaces"], rec["surfaces"]) and (ref["status"] == 0).all()
    got = run_direct(chamber, solenoid, cfg, fp, rec["surfaces"], rec["n_surfaces"], precision)
    check(got, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pooled_navigation_and_persistent_schedule(chamber, solenoid, precision):
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=True)
    fp = da.generate_tracks("toy", 4000, 50, eta_min=-1.0, eta_max=1.0)
    ref = oracle_host.propagate(chamber, solenoid, cfg, fp, threads=oracle_host.default_threads())
    order = torch.from_numpy(np.random.default_rng(5).permutation(fp.shape[1]).astype(np.int32)).cuda()
    for launch in (da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "wave"),
                   da.LaunchConfig(precision, "persistent", 256, 2, 64, order, 8, "wave")):
        check(run_launch(chamber, solenoid, cfg, fp, launch), ref, precision)


@pytest.mark.parametrize("precision", [None, "fp64", "mixed"])
def test_field_gradient_matches_oracle(chamber, solenoid, precision):
    """use_field_gradient on wires: the line's dpos/dangle block passes
    through the position columns of the transport Jacobian (None: through
    dtr_propagate)."""
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=True, use_field_gradient=True)
    fp = da.generate_tracks("toy", 4000, 51, eta_min=-1.0, eta_max=1.0)
    ref = oracle_host.propagate(chamber, solenoid, cfg, fp, threads=oracle_host.default_threads())
    if precision is None:
        got, _ = run_both(chamber, solenoid, cfg, fp)
        assert_parity(got, ref)
    else:
        got = run_launch(chamber, solenoid, cfg, fp, da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "lane"))
        check(got, ref, precision)


def test_field_gradient_rejects_pooled_navigation(chamber, solenoid):
    tracks = torch.from_numpy(da.generate_tracks("toy", 64, 47)).cuda()
    cfg = da.PropagationConfig(search_window=(3, 3), use_field_gradient=True)
    with pytest.raises(capi.DetrayError, match="UNSUPPORTED"):
        da.Propagator(cfg, da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "wave")).propagate(
            chamber, solenoid, tracks)
