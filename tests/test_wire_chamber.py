"""Wire surfaces (line_square / line_circular masks, line2D frame, material
rods, grids with several surfaces per bin) and the wire chamber in the
oracle, no GPU.

Pins:
* the reference's line intersector unit tests
  (tests/unit_tests/cpu/navigation/intersection/line_intersector.cpp:
  case1, case2, square_scope): path, signed loc0 ("left/right"), loc1,
  inside/outside;
* the wire chamber (build_wire_chamber.hpp): 11 volumes (the reference's
  JSON round trip of it, io_json_detector_roundtrip.cpp:319-342), wires on
  their layer's central radius with the generator's phi step, alternating
  stereo angle, every wire in its phi bin, tungsten rods on the wires only;
* navigation against a brute-force ray scan over every wire (the
  reference's wire_chamber_navigation straight-line check,
  tests/integration_tests/cpu/detectors/wire_chamber_navigation.cpp:48-62):
  identical wire sequences in zero field;
* the covariance transport through line frames against central finite
  differences of the propagated bound parameters: exact in zero field; in a
  2 T field within the error the navigation's path tolerance leaves (a
  track is "on" a wire within 1 um of its closest approach), and without
  the direction or dt/ds terms of the line path derivative
  (jacobian_line.hpp:64-92) the agreement is measurably worse;
* rod energy loss scales with the path 2 sqrt((r^2 - d^2) / (1 - cos^2))
  (material_rod.hpp:62-78) and vanishes for |d| > r.
"""

import math

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host
from tests.test_field_gradient import _cov_full


def transform(t=(0.0, 0.0, 0.0), z=(0.0, 0.0, 1.0), x=(1.0, 0.0, 0.0)):
    z = np.asarray(z, float) / np.linalg.norm(z)
    x = np.asarray(x, float) / np.linalg.norm(x)
    y = np.cross(z, x)
    tr = capi.Transform()
    for r in range(3):
        tr.r[r * 3 + 0], tr.r[r * 3 + 1], tr.r[r * 3 + 2] = x[r], y[r], z[r]
        tr.t[r] = t[r]
    return tr


def line_mask(shape, size, hz=math.inf):
    m = capi.Mask()
    m.v[0], m.v[1] = size, hz
    m.shape = shape
    return m


TOL = 1e-5


def test_line_intersector_case1():
    ln = line_mask(capi.LINE_CIRCULAR, 10.0)
    tf = transform()
    ok, s, l0, l1, inside = oracle_host.line_intersect((1, -1, 0), (0, 1, 0), tf, ln, TOL)
    assert ok and inside and s == 1.0 and l0 == -1.0 and l1 == 0.0  # right
    ok, s, l0, l1, inside = oracle_host.line_intersect((-1, -1, 0), (0, 1, 0), tf, ln, TOL)
    assert ok and inside and s == 1.0 and l0 == 1.0 and l1 == 0.0  # left
    d = np.array([0.0, 1.0, -1.0]) / math.sqrt(2.0)
    ok, s, l0, l1, inside = oracle_host.line_intersect((1, -1, 2), d, tf, ln, TOL)
    assert ok and inside
    assert s == pytest.approx(math.sqrt(2.0), abs=1e-5)
    assert l0 == pytest.approx(-1.0, abs=1e-5) and l1 == pytest.approx(1.0, abs=1e-5)


def test_line_intersector_case2():
    tf = transform(t=(1, 1, 1), z=(1, 0, 1), x=(1, 0, -1))
    ok, s, l0, l1, inside = oracle_host.line_intersect((1, -1, 0), (0, 1, 0), tf, line_mask(capi.LINE_CIRCULAR, 10.0),
                                                       TOL)
    assert ok and inside
    assert s == pytest.approx(2.0, abs=1e-5)
    assert l0 == pytest.approx(-1.0 / math.sqrt(2.0), abs=1e-5)
    assert l1 == pytest.approx(-1.0 / math.sqrt(2.0), abs=1e-5)


def test_line_intersector_square_scope():
    ln = line_mask(capi.LINE_SQUARE, 1.0)
    tf = transform()
    s2 = 1.0 / math.sqrt(2.0)
    cases = [((2, 0, 0), (-1, 1, 0)), ((1.9, 0, 0), (-1, 1, 0)), ((2.1, 0, 0), (-1, 1, 0)),
             ((-2, 0, 0), (1, 1, 0)), ((-1.9, 0, 0), (1, 1, 0)), ((-2.1, 0, 0), (1, 1, 0)),
             ((0, -2, 0), (-1, 1, 0)), ((0, -1.9, 0), (-1, 1, 0)), ((0, -2.1, 0), (-1, 1, 0)),
             ((0, -2, 0), (1, 1, 0)), ((0, -1.9, 0), (1, 1, 0)), ((0, -2.1, 0), (1, 1, 0))]
    res = [oracle_host.line_intersect(p, np.array(d, float) * s2, tf, ln, TOL) for p, d in cases]
    assert res[0][1] == pytest.approx(math.sqrt(2.0), abs=1e-5)
    assert res[0][2] == pytest.approx(-math.sqrt(2.0), abs=1e-5) and res[0][3] == pytest.approx(0.0, abs=1e-5)
    inside = [r[4] for r in res]
    assert inside == [True, True, False] * 4
    negative = [math.copysign(1.0, r[2]) < 0 for r in res]
    assert negative == [True] * 3 + [False] * 3 + [False] * 3 + [True] * 3


def test_parallel_ray_has_no_intersection():
    ok, *_ = oracle_host.line_intersect((1, 0, 0), (0, 0, 1), transform(), line_mask(capi.LINE_CIRCULAR, 10.0), TOL)
    assert not ok


@pytest.fixture(scope="module")
def chamber():
    return da.Detector.wire_chamber(half_z=500.0)


def test_wire_chamber_layout(chamber):
    d = chamber.desc
    assert d.n_volumes == 11 and d.n_grids == 10
    vols, sfs, trfs, masks = (chamber.table(k) for k in ("volumes", "surfaces", "transforms", "masks"))
    for v in range(1, 11):
        vol = vols[v]
        layer = [sfs[s] for s in range(vol.sf_begin, vol.sf_end)]
        assert [s.sf_id for s in layer[:4]] == [0] * 4 and all(s.sf_id == 1 for s in layer[4:])
        wires = layer[4:]
        inner = 500.0 + (v - 1) * 20.0
        central = inner + 10.0
        t = np.array([trfs[s.transform].t[:] for s in wires])
        assert np.allclose(np.hypot(t[:, 0], t[:, 1]), central, rtol=1e-6) and np.all(t[:, 2] == 0.0)
        ph = np.unwrap(np.arctan2(t[:, 1], t[:, 0]))
        delta = 20.0 / central
        assert np.allclose(np.diff(ph), delta, rtol=1e-4)
        assert ph[-1] <= 2 * math.pi + 1e-6 and ph[-1] + delta > 2 * math.pi - 1e-6
        # wire axis tilted by the stereo angle about the radial direction,
        # the sign alternating from layer to layer
        z = np.array([np.array(trfs[s.transform].r[:]).reshape(3, 3)[:, 2] for s in wires])
        assert np.allclose(z[:, 2], math.cos(0.05), rtol=1e-6)
        tilt = np.cross(t / np.linalg.norm(t, axis=1)[:, None], z)[:, 2]
        assert np.all(np.sign(tilt) == (-1.0 if (v - 1) % 2 == 0 else 1.0))
        assert all(masks[s.mask].shape == capi.LINE_SQUARE and masks[s.mask].v[0] == 10.0 and
                   masks[s.mask].v[1] == 500.0 and masks[s.mask].volume_link == v for s in wires)
        # every wire in the phi bin of its position (100 bins, up to 3 wires each)
        g = chamber.desc.grids[vol.grid]
        assert (g.n0, g.n1, g.capacity) == (100, 1, 3)
        bins = np.array(chamber.table("grid_bins")[g.bins:g.bins + 300]).reshape(100, 3)
        for k, s in enumerate(range(vol.sf_begin + 4, vol.sf_end)):
            b = int((math.atan2(t[k, 1], t[k, 0]) + math.pi) / (2 * math.pi / 100.0))
            assert s in bins[min(b, 99)]
    mats = chamber.table("materials")
    assert len(mats) == 1 and mats[0].kind == capi.MATERIAL_ROD and mats[0].thickness == 0.015
    assert mats[0].Z == 74.0
    assert all((s.material != 0xFFFFFFFF) == (s.sf_id == 1) for s in sfs)


def test_navigation_matches_ray_scan(chamber):
    fp = da.generate_tracks("toy", 2000, 42, eta_min=-0.8, eta_max=0.8, sigma_z=0.0, sigma_xy=0.0)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False, min_mask_tolerance=TOL,
                               max_mask_tolerance=TOL, max_surfaces=64)
    r = oracle_host.propagate(chamber, da.MagneticField.constant(0, 0, 0), cfg, fp)
    assert (r["status"] == 0).all()
    sfs, trfs, masks = chamber.table("surfaces"), chamber.table("transforms"), chamber.table("masks")
    sens = np.array([i for i, s in enumerate(sfs) if s.sf_id == 1])
    T = np.array([trfs[sfs[i].transform].t[:] for i in sens])
    R = np.array([np.array(trfs[sfs[i].transform].r[:]).reshape(3, 3) for i in sens])
    M = np.array([masks[sfs[i].mask].v[:2] for i in sens])
    z = R[:, :, 2]
    n_hits = 0
    for k in range(fp.shape[1]):
        ro, rd = fp[0:3, k], fp[4:7, k]
        zd = z @ rd
        t2l = T - ro
        s = (t2l @ rd - (t2l * z).sum(1) * zd) / (1 - zd * zd)
        loc = np.einsum("nji,nj->ni", R, ro + s[:, None] * rd - T)
        ok = ((s > 0) & (np.abs(loc[:, 0]) <= M[:, 0] + TOL) & (np.abs(loc[:, 1]) <= M[:, 0] + TOL) &
              (np.abs(loc[:, 2]) <= M[:, 1] + TOL))
        expect = sens[ok][np.argsort(s[ok])]
        np.testing.assert_array_equal(r["surfaces"][:r["n_surfaces"][k], k], expect, err_msg=f"track {k}")
        n_hits += len(expect)
    assert n_hits > 8 * fp.shape[1]


def _fd_check(det, field, start_sf, start, h, **kw):
    cols = 13
    bp = np.tile(np.asarray(start, float)[:, None], (1, cols))
    for j in range(6):
        bp[j, 1 + 2 * j] += h[j]
        bp[j, 2 + 2 * j] -= h[j]
    c0 = np.zeros((21, cols))
    for k in (0, 6, 11, 15, 18, 20):
        c0[k] = 1.0
    cfg = da.PropagationConfig(estimate_scattering_noise=False, search_window=(3, 3), **kw)
    out = oracle_host.propagate(det, field, cfg, None, initial_cov=c0, bound_params=bp,
                                start_surface=np.full(cols, start_sf, dtype=np.uint32))
    assert (out["n_surfaces"] == out["n_surfaces"][0]).all() and out["n_surfaces"][0] >= 5
    J = np.zeros((6, 6))
    for j in range(6):
        J[:, j] = (out["bound_params"][:, 1 + 2 * j] - out["bound_params"][:, 2 + 2 * j]) / (2 * h[j])
    cov = _cov_full(out["bound_cov"][:, :1])[:, :, 0]
    ref = J @ J.T
    d = np.sqrt(np.abs(np.diag(ref)))
    return (np.abs(cov - ref) / np.outer(d, d)).max()


@pytest.fixture(scope="module")
def bare_chamber():
    return da.Detector.wire_chamber(half_z=500.0, mat_radius=0.0)


def _starts(det):
    s0 = det.desc.volumes[1].sf_begin + 14
    t = det.desc.transforms[det.desc.surfaces[s0].transform]
    ph = math.atan2(t.t[1], t.t[0])
    return s0, [(2.0, 30.0, ph + 0.05, math.pi / 2 - 0.1, -0.2, 0.0),
                (-4.0, -100.0, ph - 0.1, math.pi / 2 + 0.3, 0.5, 0.0)]


H = np.array([1e-4, 1e-4, 1e-6, 1e-6, 1e-6, 1e-4])


def test_line_covariance_transport_zero_field(bare_chamber):
    s0, starts = _starts(bare_chamber)
    for st in starts:
        assert _fd_check(bare_chamber, da.MagneticField.constant(0, 0, 0), s0, st, H) < 1e-6


def test_line_covariance_transport_in_field(bare_chamber):
    s0, (a, b) = _starts(bare_chamber)
    field = da.MagneticField.constant(0, 0, 2 * da.TESLA)
    assert _fd_check(bare_chamber, field, s0, a, H) < 5e-4
    assert _fd_check(bare_chamber, field, s0, b, H) < 2e-5


def _energy_after_start_wire(det, start_sf, loc0, theta):
    t = det.desc.transforms[det.desc.surfaces[start_sf].transform]
    ph = math.atan2(t.t[1], t.t[0])
    bp = np.array([[loc0], [0.0], [ph], [theta], [-1.0 / 0.5], [0.0]])
    cfg = da.PropagationConfig(estimate_scattering_noise=False, search_window=(3, 3), max_surfaces=1)
    out = oracle_host.propagate(det, da.MagneticField.constant(0, 0, 0), cfg, None, bound_params=bp,
                                start_surface=np.array([start_sf], dtype=np.uint32))
    qop = out["bound_params"][4, 0]
    m = 0.1056583755
    return math.sqrt(m * m + 1.0 / (qop * qop))


def test_rod_energy_loss_follows_the_path_segment(chamber):
    s0, _ = _starts(chamber)
    m = 0.1056583755
    e0 = math.sqrt(m * m + 0.25)
    r = 0.015
    wire_z = np.array(chamber.desc.transforms[chamber.desc.surfaces[s0].transform].r[:]).reshape(3, 3)[:, 2]
    loss = {}
    for a in (0.0, 0.005, 0.010, 0.014, 0.020):
        for th in (math.pi / 2, math.pi / 2 - 0.4):
            loss[a, th] = e0 - _energy_after_start_wire(chamber, s0, a, th)
    assert loss[0.0, math.pi / 2] > 0.0
    for th in (math.pi / 2, math.pi / 2 - 0.4):
        for a in (0.005, 0.010, 0.014):
            assert loss[a, th] / loss[0.0, th] == pytest.approx(math.sqrt(1 - (a / r) ** 2), rel=1e-3)
        assert loss[0.020, th] == 0.0

    def sin_inc(th):
        t = chamber.desc.transforms[chamber.desc.surfaces[s0].transform]
        ph = math.atan2(t.t[1], t.t[0])
        d = np.array([math.cos(ph) * math.sin(th), math.sin(ph) * math.sin(th), math.cos(th)])
        return math.sqrt(1 - np.dot(d, wire_z) ** 2)

    ratio = loss[0.0, math.pi / 2 - 0.4] / loss[0.0, math.pi / 2]
    assert ratio == pytest.approx(sin_inc(math.pi / 2) / sin_inc(math.pi / 2 - 0.4), rel=1e-3)
