"""Direct navigator (direct_navigator.hpp:29-298): the oracle's restatement
checked against the reference's own integration test
(tests/integration_tests/cpu/propagator/propagation_direct_navigator.cpp:
46-230 for the toy detector, 236-420 for the wire chamber): tracks are
propagated with the caching navigator, the recorded surface sequence is
replayed forward with the direct navigator, which must visit the same
surfaces and reach the last one with the same momentum (2e-3 toy, 2e-5 wire
chamber); then from the bound parameters at the last surface the sequence is
replayed backward, which must visit the surfaces in reverse order and return
the initial momentum at the first surface (the material interactor gives the
energy back going backward) to the same tolerance."""

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


def _forward_backward(det, field, cfg, fp, tol):
    """The reference test's three propagations; returns the number of
    tracks checked."""
    ref = oracle_host.propagate(det, field, cfg, fp)
    has = ref["n_surfaces"] > 0
    fw = oracle_host.propagate_direct(det, field, cfg, fp, ref["surfaces"], ref["n_surfaces"])
    idx = np.nonzero(has)[0]
    last = ref["surfaces"][ref["n_surfaces"][idx] - 1, idx]
    # direct_backward_state(fw_updater_state.bound_params(), ...):
    # set_direction(e_backward), reset()
    bw = oracle_host.propagate_direct(det, field, cfg, None, ref["surfaces"][:, idx], ref["n_surfaces"][idx],
                                      backward=True, bound_params=fw["bound_params"][:, idx], start_surface=last,
                                      initial_cov=fw["bound_cov"][:, idx])
    assert (fw["status"][idx] == 0).all() and (bw["status"] == 0).all()
    np.testing.assert_array_equal(bw["n_surfaces"], ref["n_surfaces"][idx])
    for j, i in enumerate(idx):
        k = ref["n_surfaces"][i]
        np.testing.assert_array_equal(bw["surfaces"][:k, j], ref["surfaces"][:k, i][::-1])
    p0 = 1.0 / np.abs(fp[7, idx])
    p_fw, p_bw = momentum(fw["bound_params"][:, idx]), momentum(bw["bound_params"])
    assert np.max(np.abs(p_bw - p0) / p0) < tol
    assert (p_bw >= p_fw).all()
    # the backward pass steps against the track direction: a negative signed
    # path length (zero for a sequence of one surface)
    many = ref["n_surfaces"][idx] > 1
    assert (bw["path_length"][many] < 0.0).all() and (bw["path_length"][~many] == 0.0).all()
    return len(idx)


@pytest.mark.parametrize("overstep,bfield", CASES)
def test_oracle_direct_navigator_backward(toy7, overstep, bfield):
    field = da.MagneticField.constant(*(b * da.TESLA for b in bfield))
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False, overstep_tolerance=overstep)
    assert _forward_backward(toy7, field, cfg, uniform_tracks(), 2e-3) > 2000


@pytest.fixture(scope="module")
def chamber():
    return da.Detector.wire_chamber()


@pytest.mark.parametrize("overstep,bfield", CASES)
def test_oracle_direct_navigator_backward_wire_chamber(chamber, overstep, bfield):
    field = da.MagneticField.constant(*(b * da.TESLA for b in bfield))
    cfg = da.PropagationConfig(search_window=(5, 5), estimate_scattering_noise=False, overstep_tolerance=overstep)
    assert _forward_backward(chamber, field, cfg, uniform_tracks(), 2e-5) > 1000


def test_oracle_direct_navigator_backward_reverses_the_transport(toy7):
    """Backward replay inverts the forward transport: without material
    effects, parameters and covariance given on a module, carried forward
    along the sequence and back, return to themselves (the bound-to-bound
    map of the reversed path is the inverse map, J_bw J_fw = I), to the
    accuracy of the Runge-Kutta integration.  The sequences start at the
    first module: the beam pipe is a concentric cylinder, whose (portal)
    intersector takes the far crossing of a ray coming from outside
    (ray_cylinder_portal_intersector.hpp:68-132), so a backward pass ending
    there lands on the opposite side, as the reference's would."""
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False, do_energy_loss=False,
                               do_multiple_scattering=False)
    fp = uniform_tracks(12, 12)
    for bz in (0.0, 2.0):
        field = da.MagneticField.constant(0.0, 0.0, bz * da.TESLA)
        ref = oracle_host.propagate(toy7, field, cfg, fp)
        idx = np.nonzero(ref["n_surfaces"] > 2)[0]
        assert len(idx) > 50
        # sequences from the first module on (entry 0 is the beam pipe)
        seq = np.full_like(ref["surfaces"][:, idx], 0xFFFFFFFF)
        seq[:-1] = ref["surfaces"][1:, idx]
        nseq = ref["n_surfaces"][idx] - 1
        assert (ref["surfaces"][0, idx] == 0).all()
        first, last = seq[0], seq[nseq - 1, np.arange(len(idx))]
        start = oracle_host.propagate_direct(toy7, field, cfg, fp[:, idx], seq, np.ones_like(nseq))
        cov0 = np.zeros((21, len(idx)))
        for k, v in zip((0, 6, 11, 15, 18, 20), (1e-2, 4e-2, 1e-6, 2e-6, 1e-6, 1.0)):
            cov0[k] = v
        fw = oracle_host.propagate_direct(toy7, field, cfg, None, seq, nseq, bound_params=start["bound_params"],
                                          start_surface=first, initial_cov=cov0)
        bw = oracle_host.propagate_direct(toy7, field, cfg, None, seq, nseq, backward=True,
                                          bound_params=fw["bound_params"], start_surface=last,
                                          initial_cov=fw["bound_cov"])
        assert (fw["status"] == 0).all() and (bw["status"] == 0).all()
        # the transport did something: the forward covariance differs
        assert np.max(np.abs(fw["bound_cov"][0] - cov0[0]) / cov0[0]) > 1.0
        # a surface counts as reached within the path tolerance (1 um), so
        # each pass lands up to 1 um along the track from the exact crossing
        d = np.abs(bw["bound_params"] - start["bound_params"])
        assert d[:2].max() < 2e-3 and d[2:].max() < 1e-6, (d[:2].max(), d[2:].max())
        scale = np.sqrt(cov0[[0, 6, 11, 15, 18, 20]])
        iu = np.triu_indices(6)
        err = np.abs(bw["bound_cov"] - cov0) / (scale[iu[0]] * scale[iu[1]])
        # 1e-5 measured in zero field, 2e-5 in 2 T (Runge-Kutta accuracy and
        # the 1 um landing of each pass); a wrong sign or a dropped term of
        # the backward transport gives errors of order one
        assert err.max() < 1e-4, (bz, err.max())
