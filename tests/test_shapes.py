"""Annulus (annulus2D, polar2D frame) and placed cylinder (cylinder2D,
cylindrical2D frame, two intersections per ray) surfaces in the oracle, no
GPU.

Pins:
* the reference's intersection kernel test
  (tests/unit_tests/cpu/navigation/intersection/intersection_kernel.cpp:
  94-240): a ray through a rectangle, a trapezoid, an annulus, a rotated
  cylinder (entered and left: two candidates), a concentric portal cylinder
  and a shifted rectangle accepted only through the external tolerance
  gives the seven expected global points;
* navigation through a telescope of annuli and a telescope of cylinders
  across the beam: every cylinder is met twice (in and out), annuli cut by
  their phi range are skipped;
* the covariance transport through the polar frame (annulus) and the
  cylindrical frame (cylinder2D) against central finite differences of the
  propagated bound parameters.
"""

import ctypes as C
import math

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host
from tests.test_field_gradient import _cov_full


def trf(t=(0.0, 0.0, 0.0), z=(0.0, 0.0, 1.0), x=(1.0, 0.0, 0.0)):
    z = np.asarray(z, float)
    x = np.asarray(x, float)
    y = np.cross(z, x)
    tr = capi.Transform()
    for r in range(3):
        tr.r[r * 3 + 0], tr.r[r * 3 + 1], tr.r[r * 3 + 2] = x[r], y[r], z[r]
        tr.t[r] = t[r]
    return tr


def mask(shape, values, link=0):
    m = capi.Mask()
    for i, v in enumerate(values):
        m.v[i] = v
    m.shape = shape
    m.volume_link = link
    return m


def surface(index, trf_idx, mask_idx, shape, sf_id):
    s = capi.Surface()
    s.barcode = index
    s.transform = trf_idx
    s.mask = mask_idx
    s.material = 0xFFFFFFFF
    s.volume = 0
    s.n_masks = 1
    s.sf_id = sf_id
    s.shape = shape
    return s


def test_intersection_kernel_kat():
    shapes = [
        (trf((0, 0, 10)), mask(capi.RECTANGLE2D, (10, 10)), 1),
        (trf((0, 0, 20)), mask(capi.TRAPEZOID2D, (10, 20, 30, 1 / 60)), 1),
        (trf((0, -20, 30)), mask(capi.ANNULUS2D, (15, 55, 0.75, 1.95, 0, 2, -2)), 1),
        (trf((0, 0, 50), z=(1, 0, 0), x=(0, 0, -1)), mask(capi.CYLINDER2D, (5, -10, 10)), 2),
        (trf(), mask(capi.CONCENTRIC_CYLINDER2D, (1, 0, 1000)), 0),
        (trf((12, 0, 1000)), mask(capi.RECTANGLE2D, (10, 10)), 1),
    ]
    transforms = [capi.Transform()] + [t for t, _, _ in shapes]
    masks = [m for _, m, _ in shapes]
    surfaces = [surface(i, i + 1, i, m.shape, sid) for i, (_, m, sid) in enumerate(shapes)]
    vol = capi.Volume()
    vol.transform, vol.sf_begin, vol.sf_end, vol.bf_begin, vol.bf_end, vol.grid = 0, 0, 6, 0, 6, -1
    det = da.Detector.from_tables([vol], surfaces, transforms, masks, bf_surfaces=range(6))
    d = np.array([0.01, 0.01, 10.0])
    d /= np.linalg.norm(d)
    found = []
    for s in range(6):
        found += [c for c in oracle_host.surface_candidates(det, s, (0, 0, 0), d, 1e-3, ext_tol=1.0)]
    found.sort(key=lambda c: c[0])
    expected = [(0.01, 0.01, 10), (0.02, 0.02, 20), (0.03, 0.03, 30), (0.045, 0.045, 45), (0.055, 0.055, 55),
                (0.7071068, 0.7071068, 707.1068), (1, 1, 1000)]
    assert len(found) == len(expected)
    for (path, g, status), e in zip(found, expected):
        np.testing.assert_allclose(g, e, atol=1e-3)
        assert path > 0
    # the shifted rectangle is met on its edge only (external tolerance)
    assert [c[2] for c in found] == [0, 0, 0, 0, 0, 0, 1]


def _modified_telescope(kind, n=6, thickness=0.15):
    """The telescope (planes at 100, 200, ... mm) with its planes replaced by
    annuli or by cylinders of radius 20 mm whose axis is the x axis."""
    base = da.Detector.telescope([100.0 * i for i in range(1, n + 1)], 100.0, 100.0, thickness, 60.0)
    copy = lambda items, t: [t.from_buffer_copy(x) for x in items]
    volumes = copy(base.table("volumes"), capi.Volume)
    surfaces = copy(base.table("surfaces"), capi.Surface)
    transforms = copy(base.table("transforms"), capi.Transform)
    masks = copy(base.table("masks"), capi.Mask)
    materials = copy(base.table("materials"), capi.MaterialSlab)
    for s in surfaces:
        if s.sf_id != 1:
            continue
        m = masks[s.mask]
        z = transforms[s.transform].t[2]
        if kind == "annulus":
            # full-width ring sectors, shifted beam centre; every other plane
            # leaves out phi in (-1, 1) around its average
            k = int(round(z / 100.0))
            half = 3.1 if k % 2 else 2.0
            vals = (1.0, 150.0, -half, half, math.pi if k % 2 == 0 else 0.0, 3.0, -2.0)
            masks[s.mask] = mask(capi.ANNULUS2D, vals, m.volume_link)
            s.shape = capi.ANNULUS2D
        else:
            transforms[s.transform] = trf((0, 0, z), z=(1, 0, 0), x=(0, 0, -1))
            masks[s.mask] = mask(capi.CYLINDER2D, (20.0, -50.0, 50.0), m.volume_link)
            s.shape = capi.CYLINDER2D
    return da.Detector.from_tables(volumes, surfaces, transforms, masks, materials,
                                   bf_surfaces=base.table("bf_surfaces"), name=f"{kind}_telescope")


def _gridded_cylinder_telescope(n=3):
    """The cylinder telescope with its cylinders moved from the brute-force
    list into one ring-grid bin of capacity n (a placed cylinder in a grid
    bin, next to the others)."""
    base = _modified_telescope("cylinder", n=n)
    copy = lambda items, t: [t.from_buffer_copy(x) for x in items]
    volumes = copy(base.table("volumes"), capi.Volume)
    surfaces = copy(base.table("surfaces"), capi.Surface)
    cyl = [i for i, sf in enumerate(surfaces) if sf.shape == capi.CYLINDER2D]
    bf = [i for i in base.table("bf_surfaces") if i not in cyl]
    volumes[0].bf_begin, volumes[0].bf_end, volumes[0].grid = 0, len(bf), 0
    g = capi.Grid()
    g.type, g.capacity, g.n0, g.n1, g.bins = 1, n, 1, 1, 0
    g.min0, g.max0, g.min1, g.max1 = 0.0, 1e5, -math.pi, math.pi
    return da.Detector.from_tables(volumes, surfaces, copy(base.table("transforms"), capi.Transform),
                                   copy(base.table("masks"), capi.Mask),
                                   copy(base.table("materials"), capi.MaterialSlab), grids=[g], bf_surfaces=bf,
                                   grid_bins=cyl, name="gridded_cylinder_telescope")


def test_cylinders_in_a_grid_bin_navigate_as_in_the_list():
    """Placed cylinders found through a multi-surface grid bin give the same
    surface sequences and parameters as through the brute-force list."""
    field = da.MagneticField.constant(0.0, 0.2 * da.TESLA, 1.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(0, 0))
    fp = da.generate_tracks("beam", 500, 45, sigma_theta=0.1)
    a = oracle_host.propagate(_modified_telescope("cylinder", n=3), field, cfg, fp)
    b = oracle_host.propagate(_gridded_cylinder_telescope(3), field, cfg, fp)
    assert a["n_surfaces"].mean() > 3
    np.testing.assert_array_equal(a["surfaces"], b["surfaces"])
    np.testing.assert_array_equal(a["status"], b["status"])
    np.testing.assert_allclose(a["bound_params"], b["bound_params"], rtol=1e-12, atol=1e-12)


@pytest.fixture(scope="module")
def annulus_telescope():
    return _modified_telescope("annulus")


@pytest.fixture(scope="module")
def cylinder_telescope():
    # three cylinders: their six candidates and the exit portal fit the
    # navigator's 8-entry candidate cache
    return _modified_telescope("cylinder", n=3)


def test_cylinders_are_met_twice(cylinder_telescope):
    """Straight tracks against a ray scan: each cylinder whose mask accepts
    both crossings is recorded twice (in and out)."""
    fp = da.generate_tracks("beam", 500, 42, sigma_theta=0.05)
    tol = 1e-5
    cfg = da.PropagationConfig(estimate_scattering_noise=False, max_surfaces=32, min_mask_tolerance=tol,
                               max_mask_tolerance=tol)
    r = oracle_host.propagate(cylinder_telescope, da.MagneticField.constant(0, 0, 0), cfg, fp)
    assert (r["status"] == 0).all()
    det = cylinder_telescope
    sfs, trfs = det.table("surfaces"), det.table("transforms")
    twice = 0
    for k in range(fp.shape[1]):
        ro, rd = fp[0:3, k], fp[4:7, k]
        hits = []
        for i, s in enumerate(sfs):
            if s.sf_id != 1:
                continue
            c = np.array(trfs[s.transform].t[:])
            # axis along x, radius 20, |x| <= 50
            a = rd[1] ** 2 + rd[2] ** 2
            b = 2 * (rd[1] * (ro[1] - c[1]) + rd[2] * (ro[2] - c[2]))
            cc = (ro[1] - c[1]) ** 2 + (ro[2] - c[2]) ** 2 - 400.0
            disc = b * b - 4 * a * cc
            if disc <= 0:
                continue
            for sol in ((-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)):
                if sol > 0 and abs(ro[0] + sol * rd[0]) <= 50.0 + tol:
                    hits.append((sol, i))
        expect = [i for _, i in sorted(hits)]
        assert list(r["surfaces"][:r["n_surfaces"][k], k]) == expect, k
        twice += len(expect) == 6
    assert twice > 100


def test_annuli_follow_their_phi_range(annulus_telescope):
    fp = da.generate_tracks("beam", 500, 43, sigma_theta=0.1)
    cfg = da.PropagationConfig(estimate_scattering_noise=False, max_surfaces=32, min_mask_tolerance=1e-5,
                               max_mask_tolerance=1e-5)
    r = oracle_host.propagate(annulus_telescope, da.MagneticField.constant(0, 0, 0), cfg, fp)
    det = annulus_telescope
    sfs, trfs, masks = det.table("surfaces"), det.table("transforms"), det.table("masks")
    for k in range(fp.shape[1]):
        ro, rd = fp[0:3, k], fp[4:7, k]
        expect = []
        for i, s in enumerate(sfs):
            if s.sf_id != 1:
                continue
            m = masks[s.mask].v
            z = trfs[s.transform].t[2]
            p = ro + (z - ro[2]) / rd[2] * rd
            r_beam = math.hypot(p[0] + m[5], p[1] + m[6])
            phi_rel = math.atan2(p[1], p[0]) - m[4]
            tol = 1e-5
            if (r_beam + tol >= m[0] and r_beam <= m[1] + tol and
                    m[2] - tol / r_beam <= phi_rel <= m[3] + tol / r_beam):
                expect.append(i)
        assert list(r["surfaces"][:r["n_surfaces"][k], k]) == expect, k
    assert 0 < r["n_surfaces"].mean() < 6


def _fd(det, start_sf, start, field):
    h = np.array([1e-4, 1e-4, 1e-6, 1e-6, 1e-6, 1e-4])
    cols = 13
    bp = np.tile(np.asarray(start, float)[:, None], (1, cols))
    for j in range(6):
        bp[j, 1 + 2 * j] += h[j]
        bp[j, 2 + 2 * j] -= h[j]
    c0 = np.zeros((21, cols))
    for k in (0, 6, 11, 15, 18, 20):
        c0[k] = 1.0
    cfg = da.PropagationConfig(estimate_scattering_noise=False)
    out = oracle_host.propagate(det, field, cfg, None, initial_cov=c0, bound_params=bp,
                                start_surface=np.full(cols, start_sf, dtype=np.uint32))
    assert (out["n_surfaces"] == out["n_surfaces"][0]).all() and out["n_surfaces"][0] >= 3
    J = np.zeros((6, 6))
    for j in range(6):
        J[:, j] = (out["bound_params"][:, 1 + 2 * j] - out["bound_params"][:, 2 + 2 * j]) / (2 * h[j])
    cov = _cov_full(out["bound_cov"][:, :1])[:, :, 0]
    ref = J @ J.T
    d = np.sqrt(np.abs(np.diag(ref)))
    return (np.abs(cov - ref) / np.outer(d, d)).max()


# without material: the noise the interactor adds is not part of J C J^T
@pytest.mark.parametrize("bz", [0.0, 1.0])
def test_polar_frame_covariance_transport(bz):
    det = _modified_telescope("annulus", thickness=0.0)
    first = next(i for i, s in enumerate(det.table("surfaces")) if s.sf_id == 1)
    field = da.MagneticField.constant(0.0, 0.0, bz * da.TESLA)
    # bound start on the first annulus: r = 40 mm, phi = 0.3 (polar frame)
    e = _fd(det, first, (40.0, 0.3, 0.2, 0.15, -0.5, 0.0), field)
    assert e < 1e-6


@pytest.mark.parametrize("bz", [0.0, 1.0])
def test_cylindrical_frame_covariance_transport(bz):
    det = _modified_telescope("cylinder", n=3, thickness=0.0)
    first = next(i for i, s in enumerate(det.table("surfaces")) if s.sf_id == 1)
    field = da.MagneticField.constant(0.0, 0.0, bz * da.TESLA)
    # bound start on the first cylinder: loc0 = r phi (local phi -2.6, i.e.
    # on its upstream half), loc1 = 5 mm along the axis
    e = _fd(det, first, (20.0 * -2.6, 5.0, 0.1, 0.1, -0.5, 0.0), field)
    assert e < 1e-6
