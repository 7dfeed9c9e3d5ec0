"""Material maps on the device (material_accessor.hpp:30-48 through
pointwise_material_interactor.hpp:79-80): the plain FP64 kernel and the
batched engine (per-lane and gathered navigation, both precisions) look up
the slab of the map bin at the bound local position.

* The 4 x 1 map telescope in a field along x (tracks keep their x): every
  track equals, bit for bit, its run through the same telescope with the
  homogeneous slab of its bin, on a closed axis (edge bins clamp) and a
  circular one (bins wrap).
* The toy detector with maps on every portal and the beampipe
  (toy_det_config::use_material_maps) in the solenoid map against the
  oracle: identical surface sequences and the FP64 / mixed contracts."""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host
from tests.test_gpu_batched import check, run_launch
from tests.test_material_maps import (CFG, FIELD, XS, beam_tracks, expected_bin, homogeneous_telescope,
                                      map_telescope)

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

LAUNCHES = {
    "plain": None,
    "mixed_persistent": da.LaunchConfig(),
    "fp64_lane": da.LaunchConfig("fp64", "grid", 256, 1, 64, None, 1, "lane"),
    "mixed_gather": da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "gather", 64),
    "fp64_gather": da.LaunchConfig("fp64", "grid", 64, 2, 64, None, 1, "gather", 64),
}


@pytest.mark.parametrize("circular", [False, True])
@pytest.mark.parametrize("launch", sorted(LAUNCHES))
def test_map_bins_equal_homogeneous_slabs(circular, launch):
    fp = beam_tracks(XS)
    got = run_launch(map_telescope(circular), FIELD, CFG, fp, LAUNCHES[launch])
    bins = np.array([expected_bin(x, circular) for x in XS])
    assert (got["status"] == 0).all() and (got["n_surfaces"] >= 10).all()
    for k in range(4):
        cols = np.flatnonzero(bins == k)
        ref = run_launch(homogeneous_telescope(k), FIELD, CFG, fp, LAUNCHES[launch])
        for key in got:
            np.testing.assert_array_equal(got[key][..., cols], ref[key][..., cols], err_msg=f"bin {k}: {key}")


@pytest.fixture(scope="module")
def toy_maps():
    return da.Detector.toy(4, 7, use_material="maps")


@pytest.mark.parametrize("launch", sorted(LAUNCHES))
def test_toy_with_maps_matches_the_oracle(toy_maps, solenoid, launch):
    fp = da.generate_tracks("toy", 4096, 48)
    ref = oracle_host.propagate(toy_maps, solenoid, CFG, fp)
    got = run_launch(toy_maps, solenoid, CFG, fp, LAUNCHES[launch])
    check(got, ref, "mixed" if launch.startswith("mixed") else "fp64")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_direct_navigator_on_the_toy_with_maps(toy_maps, solenoid, precision):
    """The device direct navigator replays the oracle's sequences on the toy
    with maps and matches the oracle's replay at the batch contracts."""
    from tests.test_gpu_batched import bound_error_scale
    from tests.test_gpu_direct import run_direct
    fp = da.generate_tracks("toy", 4096, 50)
    rec = oracle_host.propagate(toy_maps, solenoid, CFG, fp)
    ref = oracle_host.propagate_direct(toy_maps, solenoid, CFG, fp, rec["surfaces"], rec["n_surfaces"])
    got = run_direct(toy_maps, solenoid, CFG, fp, rec["surfaces"], rec["n_surfaces"], precision)
    np.testing.assert_array_equal(got["status"], ref["status"])
    np.testing.assert_array_equal(got["surfaces"], ref["surfaces"])
    err = np.abs(got["bound_params"] - ref["bound_params"]) / bound_error_scale(ref)
    assert err.max() <= (1e-4 if precision == "mixed" else 1e-9), err.max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pause_and_resume_on_the_toy_with_maps(toy_maps, solenoid, precision):
    """Several calls with a step budget of 4 equal one call bit for bit."""
    from tests.test_gpu_resumable import KEYS, host, resumable
    tracks = torch.from_numpy(da.generate_tracks("toy", 4096, 51)).cuda()
    launch = da.LaunchConfig(precision, "grid", 64, 2, 64, None, 1, "gather", 64)
    ref = host(da.Propagator(CFG, launch).propagate(toy_maps, solenoid, tracks))
    got, _, first = resumable(toy_maps, solenoid, CFG, tracks, launch, 4)
    assert first > 0
    for k in KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
