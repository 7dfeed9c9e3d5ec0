"""CPU tests of the device engine's host build (tests/native/engine_host.cpp
compiles detray_amd/csrc/device/engine.hpp for the CPU):

* the FP32 grid prefilter is conservative: for rays x every grid-bin record
  of the toy detector and of both wire chambers, a surface the exact
  candidate evaluation accepts (probably inside, edge included) is never
  rejected by the prefilter.  Rays: from the beam spot and from anywhere in
  the detector, |eta| <= 3; rays aimed at points just inside the tolerance
  boundary of a mask edge (the mask tolerance pinned by min = max so the
  exact boundary is known); wires passed at distances approaching the cell
  size plus tolerance, down to rays almost parallel to the wire; external
  tolerances 0 and 5 mm.
* the per-lane propagation loop (propagate_fast, the code of every batched
  kernel) run on the host agrees with the oracle: the same surface
  sequences, bound parameters and covariances to 1e-9 (FP64 policy) and the
  same sequences with parameters to 1e-4 (mixed policy)."""

import math

import numpy as np
import pytest

import detray_amd as da
from tests import engine_host, oracle_host

RECT, TRAP, LINE_SQ, LINE_CIRC = 0, 1, 4, 5


@pytest.fixture(scope="module", params=["toy", "wire_square", "wire_round"])
def detector(request):
    if request.param == "toy":
        return request.param, da.Detector.toy(4, 7), 180.0, 1000.0
    sq = request.param == "wire_square"
    return request.param, da.Detector.wire_chamber(square_cells=sq), 600.0, 1000.0


def _directions(rng, n, eta_max):
    eta = rng.uniform(-eta_max, eta_max, n)
    phi = rng.uniform(-math.pi, math.pi, n)
    th = 2.0 * np.arctan(np.exp(-eta))
    return np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1)


def _random_rays(rng, n, r_max, z_max, beam):
    if beam:
        o = np.stack([rng.normal(0, 0.5, n), rng.normal(0, 0.5, n), rng.uniform(-100, 100, n)], axis=1)
    else:
        rr = r_max * np.sqrt(rng.uniform(0, 1, n))
        ph = rng.uniform(-math.pi, math.pi, n)
        o = np.stack([rr * np.cos(ph), rr * np.sin(ph), rng.uniform(-z_max, z_max, n)], axis=1)
    return np.concatenate([o, _directions(rng, n, 3.0)], axis=1)


def _grid_surfaces(det):
    sfs = det.table("surfaces")
    ids = sorted({int(b) for b in det.table("grid_bins") if b != 0xFFFFFFFF})
    trf = det.table("transforms")
    masks = det.table("masks")
    out = []
    for k in ids:
        s = sfs[k]
        t = trf[s.transform]
        out.append((k, s.shape, np.array(list(t.t)), np.array(list(t.r)).reshape(3, 3),
                    np.array(list(masks[s.mask].v))))
    return out


def _approach(rng, n, bound):
    """n offsets approaching `bound` from inside (1e-6 .. 1 of it) and a
    few beyond it."""
    frac = 10.0 ** rng.uniform(-6.0, 0.0, n)
    x = bound * (1.0 - frac)
    beyond = rng.uniform(0, 1, n) < 0.2
    return np.where(beyond, bound + rng.uniform(0.0, 0.05, n), x)


def _edge_rays(rng, surfaces, per_surface, tol):
    """Rays aimed at planar mask edges and passing wires near the cell
    boundary; `tol` is the total tolerance of the exact check."""
    rays = []
    for k, shape, c, R, v in surfaces:
        m = per_surface
        if shape in (RECT, TRAP):
            hy = v[1] if shape == RECT else v[2]
            y_edge = rng.uniform(0, 1, m) < 0.5
            sgn = rng.choice([-1.0, 1.0], size=(m, 2))
            y = np.where(y_edge, _approach(rng, m, hy + tol), rng.uniform(0.0, hy, m)) * sgn[:, 1]
            if shape == RECT:
                hx = np.full(m, v[0])
            else:
                hx = v[0] + (v[2] + y) * v[3] * (v[1] - v[0])
            x = np.where(y_edge, rng.uniform(0.0, 1.0, m) * hx, _approach(rng, m, hx + tol)) * sgn[:, 0]
            target = c + (R @ np.stack([x, y, np.zeros(m)])).T
            d = _directions(rng, m, 3.0)
            # keep the incidence off grazing so that the ray reaches the plane
            cosi = np.abs(d @ R[:, 2])
            d = np.where((cosi > 0.05)[:, None], d, np.tile(R[:, 2], (m, 1)))
            o = target - rng.uniform(20.0, 1500.0, m)[:, None] * d
        elif shape in (LINE_SQ, LINE_CIRC):
            half_cell, half_z = v[0], v[1]
            ang = rng.uniform(-math.pi, math.pi, m)
            if shape == LINE_SQ:
                # corners and faces of the square cell
                rad = _approach(rng, m, half_cell + tol)
                a = np.where(rng.uniform(0, 1, m) < 0.5, rad, rng.uniform(-1, 1, m) * rad)
                b = np.where(rng.uniform(0, 1, m) < 0.5, rad, rng.uniform(-1, 1, m) * rad)
                sw = rng.uniform(0, 1, m) < 0.5
                a, b = np.where(sw, a, b), np.where(sw, b, a)
                a *= rng.choice([-1.0, 1.0], m)
                b *= rng.choice([-1.0, 1.0], m)
            else:
                rad = _approach(rng, m, half_cell + tol)
                a, b = rad * np.cos(ang), rad * np.sin(ang)
            z = rng.uniform(-half_z, half_z, m) * 0.98
            target = c + (R @ np.stack([a, b, z])).T
            # direction in the plane of the wire axis and the tangent at the
            # target, so the target is the point of closest approach; angles
            # to the axis from 90 degrees down to almost parallel
            radial = np.stack([a, b, np.zeros(m)], axis=1)
            radial /= np.linalg.norm(radial, axis=1, keepdims=True)
            tang_loc = np.cross(np.array([0.0, 0.0, 1.0]), radial)
            alpha = np.where(rng.uniform(0, 1, m) < 0.3, rng.uniform(1.45, 1.5707, m), rng.uniform(0.0, 1.45, m))
            d_loc = np.cos(alpha)[:, None] * tang_loc + (np.sin(alpha) * rng.choice([-1.0, 1.0], m))[:, None] * \
                np.array([0.0, 0.0, 1.0])
            d = (R @ d_loc.T).T
            o = target - rng.uniform(20.0, 1500.0, m)[:, None] * d
        else:
            continue
        rays.append(np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], axis=1))
    return np.concatenate(rays, axis=0)


@pytest.mark.parametrize("ext_tol", [0.0, 5.0])
def test_prefilter_is_conservative_on_random_rays(detector, ext_tol):
    name, det, r_max, z_max = detector
    rng = np.random.default_rng(7)
    cfg = da.PropagationConfig()
    for beam in (True, False):
        rays = _random_rays(rng, 1500, r_max, z_max, beam)
        n, acc, miss, first = engine_host.prefilter(det, cfg, rays, ext_tol)
        assert acc > 500, (name, beam, acc)
        assert miss == 0, (name, beam, ext_tol, first)


@pytest.mark.parametrize("tol,ext_tol", [(1e-5, 0.0), (3.0, 0.0), (0.5, 5.0)])
def test_prefilter_is_conservative_at_the_mask_edges(detector, tol, ext_tol):
    name, det, _, _ = detector
    rng = np.random.default_rng(int(tol * 1000) + int(ext_tol))
    # min = max: the exact check's tolerance is tol + ext_tol at any path
    cfg = da.PropagationConfig(min_mask_tolerance=tol, max_mask_tolerance=tol)
    surfaces = _grid_surfaces(det)
    step = max(1, len(surfaces) // 400)
    rays = _edge_rays(rng, surfaces[::step], 12, tol + ext_tol)
    n, acc, miss, first = engine_host.prefilter(det, cfg, rays, ext_tol)
    # most rays hit their own surface inside the tolerance
    assert acc > 0.5 * rays.shape[0], (name, acc, rays.shape[0])
    assert miss == 0, (name, tol, ext_tol, first)


# ------------------------------------------------- host lane loop vs oracle --
@pytest.fixture(scope="module")
def toy_run():
    det = da.Detector.toy(4, 7)
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 300, 17, pt_min=0.5, pt_max=50.0)
    n = fp.shape[1]
    cov = np.zeros((21, n))
    for k, v in {0: 1e-3, 6: 1e-3, 11: 1e-6, 15: 1e-6, 18: 1e-6, 20: 1.0}.items():
        cov[k] = v
    ref = oracle_host.propagate(det, field, cfg, fp, initial_cov=cov)
    return det, field, cfg, fp, cov, ref


def _same_sequences(got, ref):
    assert np.array_equal(got["n_surfaces"], ref["n_surfaces"])
    assert np.array_equal(got["status"], ref["status"])
    for i in range(ref["n_surfaces"].shape[0]):
        k = int(ref["n_surfaces"][i])
        assert np.array_equal(got["surfaces"][:k, i], ref["surfaces"][:k, i]), i


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)))


def test_host_lane_loop_fp64_matches_the_oracle(toy_run):
    det, field, cfg, fp, cov, ref = toy_run
    got = engine_host.propagate(det, field, cfg, fp, initial_cov=cov)
    _same_sequences(got, ref)
    assert np.array_equal(got["n_steps"], ref["n_steps"])
    assert _rel(got["bound_params"], ref["bound_params"]) < 1e-9
    assert _rel(got["path_length"], ref["path_length"]) < 1e-9
    scale = np.maximum(np.abs(ref["bound_cov"]).max(axis=0), 1e-30)
    assert float(np.max(np.abs(got["bound_cov"] - ref["bound_cov"]) / scale)) < 1e-8


def test_host_lane_loop_mixed_matches_the_oracle(toy_run):
    det, field, cfg, fp, cov, ref = toy_run
    got = engine_host.propagate(det, field, cfg, fp, mixed=True, initial_cov=cov)
    _same_sequences(got, ref)
    assert _rel(got["bound_params"], ref["bound_params"]) < 1e-4
    scale = np.maximum(np.abs(ref["bound_cov"]).max(axis=0), 1e-30)
    assert float(np.max(np.abs(got["bound_cov"] - ref["bound_cov"]) / scale)) < 1e-3


def test_host_lane_loop_wire_chamber_matches_the_oracle():
    det = da.Detector.wire_chamber()
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 200, 5, eta_min=-1.0, eta_max=1.0, pt_min=1.0, pt_max=20.0)
    ref = oracle_host.propagate(det, field, cfg, fp)
    got = engine_host.propagate(det, field, cfg, fp)
    _same_sequences(got, ref)
    assert _rel(got["bound_params"], ref["bound_params"]) < 1e-9


# ------------------------------------------- volume material, host lane loop --
def volume_material_case(name="cesium_iodide_with_ded", n=64, seed=3, **cfg_kw):
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0, volume_material=name)
    field = da.MagneticField.constant(1.0 * da.TESLA, 0.0, 0.5 * da.TESLA)
    cfg = da.PropagationConfig(**cfg_kw)
    fp = da.generate_tracks("beam", n, seed, p_min=1.0, p_max=10.0, sigma_theta=0.1)
    cov = np.zeros((21, n))
    for k, v in {0: 1e-3, 6: 1e-3, 11: 1e-6, 15: 1e-6, 18: 1e-6, 20: 1.0}.items():
        cov[k] = v
    return det, field, cfg, fp, cov


@pytest.mark.parametrize("kw", [{}, {"use_eloss_gradient": True}, {"use_mean_loss": False},
                                {"use_eloss_gradient": True, "estimate_scattering_noise": False}],
                         ids=["default", "eloss_gradient", "no_mean_loss", "gradient_no_noise"])
def test_host_lane_loop_volume_material_matches_the_oracle(kw):
    det, field, cfg, fp, cov = volume_material_case(**kw)
    ref = oracle_host.propagate(det, field, cfg, fp, initial_cov=cov)
    assert np.all(ref["n_surfaces"] == 10)
    got = engine_host.propagate(det, field, cfg, fp, initial_cov=cov)
    _same_sequences(got, ref)
    assert np.array_equal(got["n_steps"], ref["n_steps"])
    assert _rel(got["bound_params"], ref["bound_params"]) < 1e-9
    scale = np.maximum(np.abs(ref["bound_cov"]).max(axis=0), 1e-30)
    assert float(np.max(np.abs(got["bound_cov"] - ref["bound_cov"]) / scale)) < 1e-8
    mixed = engine_host.propagate(det, field, cfg, fp, mixed=True, initial_cov=cov)
    _same_sequences(mixed, ref)
    assert _rel(mixed["bound_params"], ref["bound_params"]) < 1e-4
    assert float(np.max(np.abs(mixed["bound_cov"] - ref["bound_cov"]) / scale)) < 1e-3


# ------------------------------------- direct navigator, forward and backward --
def _backward_case(det, field, cfg, fp, cov):
    """Oracle caching run, then the backward replay's inputs: bound start on
    the last recorded surface with the forward direct pass's parameters."""
    ref = oracle_host.propagate(det, field, cfg, fp, initial_cov=cov)
    idx = np.nonzero(ref["n_surfaces"] > 0)[0]
    seq, nseq = ref["surfaces"][:, idx], ref["n_surfaces"][idx]
    fw = oracle_host.propagate_direct(det, field, cfg, fp[:, idx], seq, nseq)
    last = seq[nseq - 1, np.arange(len(idx))]
    return seq, nseq, fw, last


@pytest.mark.parametrize("which", ["toy", "wire_chamber", "volume_material"])
def test_host_direct_backward_matches_the_oracle(which):
    if which == "toy":
        det = da.Detector.toy(4, 7)
        field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 1.0 * da.TESLA)
        cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
        fp = da.generate_tracks("toy", 300, 21, pt_min=0.5, pt_max=50.0)
    elif which == "wire_chamber":
        det = da.Detector.wire_chamber()
        field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
        cfg = da.PropagationConfig(search_window=(5, 5), estimate_scattering_noise=False)
        fp = da.generate_tracks("toy", 200, 22, eta_min=-1.0, eta_max=1.0, pt_min=1.0, pt_max=20.0)
    else:
        det, field, cfg, fp, _ = volume_material_case()
    seq, nseq, fw, last = _backward_case(det, field, cfg, fp, None)
    ref = oracle_host.propagate_direct(det, field, cfg, None, seq, nseq, backward=True,
                                       bound_params=fw["bound_params"], start_surface=last,
                                       initial_cov=fw["bound_cov"])
    got = engine_host.propagate_direct(det, field, cfg, None, seq, nseq, backward=True,
                                       bound_params=fw["bound_params"], start_surface=last,
                                       initial_cov=fw["bound_cov"])
    assert (ref["status"] == 0).all()
    _same_sequences(got, ref)
    assert np.array_equal(got["n_steps"], ref["n_steps"])
    assert _rel(got["bound_params"], ref["bound_params"]) < 1e-9
    assert _rel(got["path_length"], ref["path_length"]) < 1e-9
    scale = np.maximum(np.abs(ref["bound_cov"]).max(axis=0), 1e-30)
    assert float(np.max(np.abs(got["bound_cov"] - ref["bound_cov"]) / scale)) < 1e-8
    mixed = engine_host.propagate_direct(det, field, cfg, None, seq, nseq, mixed=True, backward=True,
                                         bound_params=fw["bound_params"], start_surface=last,
                                         initial_cov=fw["bound_cov"])
    _same_sequences(mixed, ref)
    assert _rel(mixed["bound_params"], ref["bound_params"]) < 1e-4
    # forward through the engine's direct loop as well
    fwd = engine_host.propagate_direct(det, field, cfg, None, seq, nseq, bound_params=ref["bound_params"],
                                       start_surface=seq[0], initial_cov=ref["bound_cov"])
    fref = oracle_host.propagate_direct(det, field, cfg, None, seq, nseq, bound_params=ref["bound_params"],
                                        start_surface=seq[0], initial_cov=ref["bound_cov"])
    _same_sequences(fwd, fref)
    assert _rel(fwd["bound_params"], fref["bound_params"]) < 1e-9
