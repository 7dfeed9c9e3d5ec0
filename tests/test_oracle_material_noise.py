"""How the material interaction and the noise estimation enter the oracle's
track state and covariance, pinned against what the physics and the
reference's definitions fix (CPU):

* energy loss through a telescope of eleven 1.7 mm silicon planes, 10 GeV
  muons: the final energy is the initial one minus the Bethe-Bloch loss of
  every plane (the recipe of
  tests/integration_tests/cpu/material/material_interaction.cpp:54-173);
* the q/p variance: the sum of the Landau sigma(q/p)^2 of every plane at
  the momentum the track had there (C(q/p, q/p) is transported with
  J(q/p, q/p) = 1 in vacuum);
* multiple scattering: the increments of C(phi, phi) and C(theta, theta) at
  a thick plane against the variance of (phi, theta) of directions
  scattered by two independent projected angles of the Highland theta0 (a
  Monte-Carlo of the scattering model, as material_interaction.cpp:175-277
  does with its random scatterer);
* noise estimation (noise_estimation.hpp:23-107): the external mask
  tolerance it sets after a material surface decides whether the next
  plane, missed by d, is still navigated to (an edge candidate); tracks
  placed at d = (1 -+ 0.3 %) of the tolerance written from the reference's
  definition are and are not recorded.

Highland's theta0 and the Bethe-Bloch and Landau expressions are pinned by
the PDG values in test_oracle_physics.py; here they are inputs."""

import math

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

MUON_MASS = 0.1056583755


def _tilted_tracks(n, p, theta, phi, x0=0.0):
    fp = np.zeros((8, n))
    fp[0] = x0
    fp[4] = math.sin(theta) * math.cos(phi)
    fp[5] = math.sin(theta) * math.sin(phi)
    fp[6] = math.cos(theta)
    fp[7] = -1.0 / p
    return fp


def _slab(det):
    return det.table("materials")[0]


def test_energy_loss_through_eleven_planes():
    thickness = 1.7
    det = da.Detector.telescope([10.0 + 50.0 * k for k in range(11)], 200.0, 200.0, thickness, 60.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    cfg = da.PropagationConfig(estimate_scattering_noise=False, do_multiple_scattering=False)
    theta = 0.02
    fp = _tilted_tracks(4, 10.0, theta, 0.4)
    res = oracle_host.propagate(det, field, cfg, fp)
    assert np.all(res["n_surfaces"] == 11)
    path = thickness / math.cos(theta)
    dedx = oracle_host.material_probe(_slab(det), fp[7, 0], path, path / _slab(det).X0)[0]
    e0 = math.hypot(10.0, MUON_MASS)
    p_new = np.abs(1.0 / res["bound_params"][4])
    e_new = np.hypot(p_new, MUON_MASS)
    # the reference holds 1e-5 GeV with the loss at the initial momentum
    assert np.max(np.abs(e_new - (e0 - 11 * dedx * path))) < 1e-5


def test_qop_variance_is_the_sum_of_the_landau_terms():
    thickness = 1.7
    det = da.Detector.telescope([10.0 + 50.0 * k for k in range(11)], 200.0, 200.0, thickness, 60.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    cfg = da.PropagationConfig(estimate_scattering_noise=False, do_multiple_scattering=False)
    theta = 0.02
    fp = _tilted_tracks(3, 2.0, theta, -1.1)
    n = fp.shape[1]
    c0 = np.zeros((21, n))
    for k, v in {0: 1e-4, 6: 1e-4, 11: 1e-6, 15: 1e-6, 18: 0.0, 20: 1.0}.items():
        c0[k] = v
    res = oracle_host.propagate(det, field, cfg, fp, initial_cov=c0, actors=da.Actors(step_tracer=16))
    st = res["steps"]
    path = thickness / math.cos(theta)
    slab = _slab(det)
    for i in range(n):
        qop_before = np.concatenate([[fp[7, i]], st.bound_params[4, :10, i]])
        sig = np.array([oracle_host.material_probe(slab, q, path, path / slab.X0)[1] for q in qop_before])
        expected = float(np.sum(sig ** 2))
        assert res["bound_cov"][18, i] == pytest.approx(expected, rel=1e-9)


@pytest.mark.parametrize("theta,phi", [(0.35, 0.7), (1.1, -2.5)])
def test_scattering_variance_against_monte_carlo(theta, phi):
    thickness = 20.0
    det = da.Detector.telescope([100.0], 800.0, 800.0, thickness, 60.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    cfg = da.PropagationConfig(estimate_scattering_noise=False, do_energy_loss=False)
    p = 3.0
    fp = _tilted_tracks(1, p, theta, phi)
    c0 = np.zeros((21, 1))
    c0[0] = c0[6] = 1e-4
    c0[18] = 1e-6
    c0[20] = 1.0
    res = oracle_host.propagate(det, field, cfg, fp, initial_cov=c0)
    d = fp[4:7, 0]
    cos_inc = abs(d[2])  # plane normal along z
    slab = _slab(det)
    theta0 = oracle_host.material_probe(slab, fp[7, 0], thickness / cos_inc, thickness / cos_inc / slab.X0)[2]
    # scattered directions: two independent projected angles of sigma theta0
    # about axes perpendicular to the direction
    rng = np.random.default_rng(21)
    m = 400_000
    e1 = np.cross(d, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    a, b = rng.normal(0.0, theta0, m), rng.normal(0.0, theta0, m)
    nd = np.cos(np.hypot(a, b))[:, None] * d + np.sinc(np.hypot(a, b) / np.pi)[:, None] * (
        a[:, None] * e1 + b[:, None] * e2)
    ph = np.arctan2(nd[:, 1], nd[:, 0])
    th = np.arctan2(np.hypot(nd[:, 0], nd[:, 1]), nd[:, 2])
    var_phi = np.mean((ph - phi) ** 2)
    var_theta = np.mean((th - theta) ** 2)
    # Monte-Carlo precision with 4e5 samples: ~0.2 %
    assert res["bound_cov"][11, 0] == pytest.approx(var_phi, rel=1e-2)
    assert res["bound_cov"][15, 0] == pytest.approx(var_theta, rel=1e-2)


def _ext_tol_from_definition(cov, phi, theta, d, path_next, p, path_length, n_std, acc_err):
    """noise_estimation.hpp:23-107 written from the reference's definition."""
    dphi = n_std * math.sqrt(cov["phi"])
    dth = n_std * math.sqrt(cov["theta"])
    pe, te = phi + dphi, theta + dth
    displ = np.array([math.cos(pe) * math.sin(te), math.sin(pe) * math.sin(te), math.cos(te)]) - d
    displ *= path_next
    acc = 1.0 / p * acc_err * path_length
    t = math.sqrt(n_std ** 2 * (cov["loc0"] + cov["loc1"]) + displ @ displ + acc * acc)
    return min(t, 5.0)


@pytest.mark.parametrize("side,expect", [(-1.0, True), (1.0, False)])
def test_noise_estimation_tolerance_decides_the_next_plane(side, expect):
    # two planes; the tracks cross plane 1 (z = 100) and pass outside the
    # +x edge of plane 2 (z = 200, half length 50) by d
    half = 50.0
    det = da.Detector.telescope([100.0, 200.0], half, 400.0, 0.3, 60.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    cfg = da.PropagationConfig(estimate_scattering_noise=True, do_multiple_scattering=False,
                               accumulated_error=0.01, n_scattering_stddev=3)
    theta, phi, p = 0.05, 0.3, 10.0
    cov = {"loc0": 0.05 ** 2, "loc1": 0.03 ** 2, "phi": 1.5e-3 ** 2, "theta": 0.8e-3 ** 2}
    c0 = np.zeros((21, 1))
    c0[0], c0[6], c0[11], c0[15], c0[18], c0[20] = cov["loc0"], cov["loc1"], cov["phi"], cov["theta"], 1e-8, 1.0
    d = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
    # a first run fixes the momentum after plane 1's energy loss and the
    # path length there (the tolerance is set after the material update)
    probe = oracle_host.propagate(det, field, cfg, _tilted_tracks(1, p, theta, phi), initial_cov=c0,
                                  actors=da.Actors(step_tracer=4))
    p1 = abs(1.0 / probe["steps"].bound_params[4, 0, 0])
    s1 = float(probe["steps"].path_length[0, 0])
    path_next = 100.0 / math.cos(theta)
    tol = _ext_tol_from_definition(cov, phi, theta, d, path_next, p1, s1, 3, 0.01)
    assert 0.2 < tol < 1.0
    # x at plane 2 = half + miss distance, the start shifted in x only
    miss = tol * (1.0 + side * 0.003)
    x0 = half + miss - 200.0 * d[0] / d[2]
    fp = _tilted_tracks(1, p, theta, phi, x0=x0)
    res = oracle_host.propagate(det, field, cfg, fp, initial_cov=c0)
    on_plane2 = int(res["n_surfaces"][0]) == 2
    assert on_plane2 == expect, (miss, tol, res["n_surfaces"][0])
    # without the noise estimation plane 2 is never reached
    cfg_off = da.PropagationConfig(estimate_scattering_noise=False, do_multiple_scattering=False)
    assert int(oracle_host.propagate(det, field, cfg_off, fp, initial_cov=c0)["n_surfaces"][0]) == 1
