"""Pins of oracle parts that no reference fixture covers, each against what
the mathematics fixes:

* trilinear field interpolation: exact for multilinear fields (any
  a + b x + c y + d z + e xy + f yz + g zx + h xyz per component) at random
  points between the nodes, and the clamp to the grid's border on every
  face and corner (the covfie affine -> linear -> clamp -> strided lookup of
  tests/include/detray/test/common/bfield.hpp:44-47);
* helical navigation through the 7-endcap toy in a constant field without
  material: every recorded surface lies on the analytic helix at its path
  length (the helix-inspector recipe of
  tests/integration_tests/cpu/propagator/propagation_caching_navigator.cpp:
  49-120, with a tighter tolerance), and an independent scan of the helix
  against every sensitive surface of the detector finds exactly the
  recorded sequence, in path order.
"""


import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

RECT, TRAP = 0, 1


# ------------------------------------------------------------- trilinear --
def _multilinear(coef, x, y, z):
    a, b, c, d, e, f, g, h = coef
    return a + b * x + c * y + d * z + e * x * y + f * y * z + g * z * x + h * x * y * z


COEFS = [
    (0.3, 1.1e-3, -0.7e-3, 0.2e-3, 2e-6, -1e-6, 0.5e-6, 3e-9),
    (-0.1, -0.4e-3, 0.9e-3, 0.6e-3, -1e-6, 2.5e-6, -0.7e-6, -2e-9),
    (1.2, 0.1e-3, 0.2e-3, -1.3e-3, 0.4e-6, 0.3e-6, 1.5e-6, 1e-9),
]


@pytest.fixture(scope="module")
def multilinear_field():
    # unequal pitches and extents, grid origin off zero
    mn = np.array([-310.0, -125.0, -640.0])
    pitch = np.array([35.0, 20.0, 80.0])
    n = np.array([19, 14, 17])
    xs, ys, zs = (mn[i] + pitch[i] * np.arange(n[i]) for i in range(3))
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    v = np.stack([_multilinear(c, X, Y, Z) for c in COEFS], axis=-1)
    field = da.MagneticField.grid(v, mn, pitch, name="multilinear")
    return field, mn, pitch, n, v.astype(np.float32)


def test_trilinear_is_exact_for_multilinear_fields(multilinear_field):
    field, mn, pitch, n, v32 = multilinear_field
    rng = np.random.default_rng(5)
    hi = mn + pitch * (n - 1)
    pts = rng.uniform(mn, hi, size=(4000, 3))
    got = oracle_host.field_at(field, pts)
    ref = np.stack([_multilinear(c, pts[:, 0], pts[:, 1], pts[:, 2]) for c in COEFS], axis=1)
    # only the fp32 rounding of the stored samples remains
    scale = np.abs(v32).max(axis=(0, 1, 2))
    assert np.max(np.abs(got - ref) / scale) < 3e-7
    # a swapped weight (t <-> 1 - t on any axis) would be off by a lot
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_trilinear_clamps_on_every_face(multilinear_field, axis):
    field, mn, pitch, n, _ = multilinear_field
    rng = np.random.default_rng(11 + axis)
    hi = mn + pitch * (n - 1)
    pts = rng.uniform(mn, hi, size=(400, 3))
    below = pts.copy()
    below[:, axis] = mn[axis] - rng.uniform(1.0, 500.0, 400)
    above = pts.copy()
    above[:, axis] = hi[axis] + rng.uniform(1.0, 500.0, 400)
    for outside in (below, above):
        clamped = np.clip(outside, mn, hi)
        got = oracle_host.field_at(field, outside)
        ref = np.stack([_multilinear(c, clamped[:, 0], clamped[:, 1], clamped[:, 2]) for c in COEFS], axis=1)
        assert np.max(np.abs(got - ref)) < 1e-6


def test_trilinear_clamps_at_the_corners(multilinear_field):
    field, mn, pitch, n, v32 = multilinear_field
    hi = mn + pitch * (n - 1)
    corners = np.array([[x, y, z] for x in (mn[0] - 90, hi[0] + 90) for y in (mn[1] - 7, hi[1] + 700)
                        for z in (mn[2] - 1000, hi[2] + 3)])
    got = oracle_host.field_at(field, corners)
    ix = np.where(corners[:, 0] < mn[0], 0, n[0] - 1)
    iy = np.where(corners[:, 1] < mn[1], 0, n[1] - 1)
    iz = np.where(corners[:, 2] < mn[2], 0, n[2] - 1)
    assert np.array_equal(got, v32[ix, iy, iz].astype(np.float64))


# --------------------------------------------------------------- helices --
def _helix(p0, t0, qop, bz):
    """Position along the helix dT/ds = qop (T x B), B = bz e_z."""
    om = qop * bz
    tx, ty, tz = t0

    def r(s):
        c, sn = np.cos(om * s), np.sin(om * s)
        return np.stack([p0[0] + (tx * sn + ty * (1.0 - c)) / om,
                         p0[1] + (ty * sn - tx * (1.0 - c)) / om,
                         p0[2] + tz * s], axis=-1)

    def t(s):
        c, sn = np.cos(om * s), np.sin(om * s)
        return np.stack([tx * c + ty * sn, ty * c - tx * sn, np.full_like(s, tz)], axis=-1)

    return r, t


@pytest.fixture(scope="module", params=[(1.0, 100.0), (0.3, 1.0)], ids=["pT1-100", "pT0.3-1"])
def helix_run(request):
    det = da.Detector.toy(4, 7, use_material=False)
    bz = 2.0 * da.TESLA
    field = da.MagneticField.constant(0.0, 0.0, bz)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = da.generate_tracks("toy", 400, 43, pt_min=request.param[0], pt_max=request.param[1])
    res = oracle_host.propagate(det, field, cfg, fp, actors=da.Actors(step_tracer=64))
    return det, bz, fp, res


def test_recorded_surfaces_lie_on_the_helix(helix_run):
    det, bz, fp, res = helix_run
    st = res["steps"]
    worst = 0.0
    for i in range(fp.shape[1]):
        m = int(st.n[i])
        if m == 0:
            continue
        r, _ = _helix(fp[0:3, i], fp[4:7, i], fp[7, i], bz)
        s = st.path_length[:m, i]
        d = np.linalg.norm(st.free_params[0:3, :m, i].T - r(s), axis=1) / s
        worst = max(worst, float(d.max()))
    # the reference's inspector holds 1e-3; the adaptive RK4 with a 1e-4 mm
    # error tolerance per step stays far inside that
    assert worst < 1e-5, worst


def _sensitive_planes(det):
    sfs = det.table("surfaces")
    trf = det.table("transforms")
    masks = det.table("masks")
    idx, c, R, shape, mv = [], [], [], [], []
    for k, s in enumerate(sfs):
        if s.sf_id != 1 or s.shape not in (RECT, TRAP):
            continue
        t = trf[s.transform]
        idx.append(k)
        c.append(list(t.t))
        R.append(np.array(t.r).reshape(3, 3))
        shape.append(s.shape)
        mv.append(list(masks[s.mask].v)[:4])
    return np.array(idx), np.array(c), np.array(R), np.array(shape), np.array(mv)


def test_helix_scan_finds_the_recorded_sequence(helix_run):
    det, bz, fp, res = helix_run
    idx, c, R, shape, mv = _sensitive_planes(det)
    normal = R[:, :, 2]
    margin = 0.05  # mm: crossings this close to a mask edge are left out
    n_checked = n_surfaces = 0
    for i in range(fp.shape[1]):
        r, t = _helix(fp[0:3, i], fp[4:7, i], fp[7, i], bz)
        s_end = float(res["path_length"][i])
        # Newton from the straight-line crossing, every plane at once
        d0 = fp[4:7, i]
        den = normal @ d0
        ok = np.abs(den) > 1e-9
        s = np.where(ok, np.einsum("kj,kj->k", normal, c - fp[0:3, i]) / np.where(ok, den, 1.0), -1.0)
        for _ in range(8):
            g = np.einsum("kj,kj->k", normal, r(s) - c)
            gp = np.einsum("kj,kj->k", normal, t(s))
            s = s - np.where(np.abs(gp) > 1e-12, g / np.where(np.abs(gp) > 1e-12, gp, 1.0), 0.0)
        resid = np.abs(np.einsum("kj,kj->k", normal, r(s) - c))
        graze = np.abs(np.einsum("kj,kj->k", normal, t(s))) < 0.05
        cand = ok & (resid < 1e-6) & (s > 1e-3) & (s < s_end - 1e-3)
        loc = np.einsum("kji,kj->ki", R, r(s) - c)
        x, y = np.abs(loc[:, 0]), loc[:, 1]
        hx = np.where(shape == TRAP, mv[:, 0] + (mv[:, 2] + y) * mv[:, 3] * (mv[:, 1] - mv[:, 0]), mv[:, 0])
        hy = np.where(shape == TRAP, mv[:, 2], mv[:, 1])
        inside = (x <= hx - margin) & (np.abs(y) <= hy - margin)
        outside = (x > hx + margin) | (np.abs(y) > hy + margin)
        ambiguous = cand & (graze | ~(inside | outside))
        expect = idx[cand & inside & ~graze]
        order = np.argsort(s[cand & inside & ~graze])
        expect = expect[order]
        m = int(res["n_surfaces"][i])
        got = res["surfaces"][:m, i].astype(np.int64)
        amb = set(idx[ambiguous].tolist())
        got_clear = [g for g in got.tolist() if g not in amb]
        assert got_clear == expect.tolist(), (i, got.tolist(), expect.tolist())
        n_checked += 1
        n_surfaces += len(expect)
    assert n_checked == fp.shape[1] and n_surfaces > 5 * n_checked
