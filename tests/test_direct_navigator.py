"""Direct navigator (direct_navigator.hpp:29-298): the oracle's restatement
checked against the reference's own integration test
(tests/integration_tests/cpu/propagator/propagation_direct_navigator.cpp:46-213,
forward part): tracks are propagated through the toy detector with the
caching navigator, the recorded surface sequence is replayed with the direct
navigator, which must visit the same surfaces and reach the last one with
the same momentum to 2e-3."""

import math

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

# (overstep tolerance, field) of the reference test's parameter sets
# (propagation_direct_navigator.cpp:216-230)
CASES = [(-0.1, (0.0, 0.0, 2.0)), (-0.4, (0.0, 1.0, 1.0)), (-0.4, (1.0, 0.0, 1.0)), (-0.6, (1.0, 1.0, 1.0))]


def uniform_tracks(n_phi=50, n_theta=50, p=10.0, q=-1.0):
    """Directions on a phi x theta grid from the origin, |p| = 10 GeV
    (uniform_track_generator with phi_steps = theta_steps = 50)."""
    fp = []
    for th in np.linspace(0.01, math.pi - 0.01, n_theta):
        for ph in np.linspace(-math.pi, math.pi, n_phi, endpoint=False):
            fp.append([0.0, 0.0, 0.0, 0.0, math.cos(ph) * math.sin(th), math.sin(ph) * math.sin(th), math.cos(th),
                       q / p])
    return np.ascontiguousarray(np.array(fp).T)


def momentum(bound):
    with np.errstate(divide="ignore"):
        return 1.0 / np.abs(bound[4])


@pytest.mark.parametrize("overstep,bfield", CASES)
def test_oracle_direct_navigator_replays_the_sequence(toy7, overstep, bfield):
    field = da.MagneticField.constant(*(b * da.TESLA for b in bfield))
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False, overstep_tolerance=overstep)
    fp = uniform_tracks()
    ref = oracle_host.propagate(toy7, field, cfg, fp)
    got = oracle_host.propagate_direct(toy7, field, cfg, fp, ref["surfaces"], ref["n_surfaces"])
    has = ref["n_surfaces"] > 0
    assert has.sum() > 2000
    assert (got["status"] == 0).all()
    np.testing.assert_array_equal(got["n_surfaces"], ref["n_surfaces"])
    for i in np.nonzero(has)[0]:
        k = ref["n_surfaces"][i]
        np.testing.assert_array_equal(got["surfaces"][:k, i], ref["surfaces"][:k, i])
    p_ref, p_dir = momentum(ref["bound_params"][:, has]), momentum(got["bound_params"][:, has])
    # propagation_direct_navigator.cpp:180-185: momentum at the last surface
    assert (p_ref < 10.0).all()
    assert np.max(np.abs(p_ref - p_dir) / p_ref) < 2e-3


def test_oracle_direct_navigator_edge_cases(toy7):
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = uniform_tracks(4, 4)
    ref = oracle_host.propagate(toy7, field, cfg, fp)
    seq = ref["surfaces"].copy()
    n_seq = ref["n_surfaces"].copy()
    n_seq[0] = 0                      # empty sequence: exit at once
    seq[1, 1] = toy7.n_surfaces + 5   # not a surface of the detector: abort
    n_seq[1] = max(n_seq[1], 2)
    got = oracle_host.propagate_direct(toy7, field, cfg, fp, seq, n_seq)
    assert got["status"][0] == 0 and got["n_surfaces"][0] == 0
    assert got["status"][1] == 1
