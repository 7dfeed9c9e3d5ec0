"""Pins of the oracle against closed forms and the reference's own
known-answer tests (not against a retyped copy of its formulas):

* Bethe-Bloch stopping power of muons in Si and Al against the PDG values the
  reference tests hold (tests/unit_tests/cpu/material/bethe_equation.cpp:
  120-160, 5 % tolerance) and the Landau FWHM KAT (bethe_equation.cpp:
  335-402, 10 %);
* Highland multiple-scattering angle against the PDG (RPP eq. 34.15) form
  written with beta from E = sqrt(p^2 + m^2);
* RK4 stepping in a constant field against the analytic helix (the recipe of
  tests/unit_tests/cpu/propagator/rk_stepper.cpp:52-145, tol 1e-3);
* the transported covariance against the Jacobian of the propagation map
  obtained by central finite differences of the oracle itself
  (jacobian_validation.cpp:139-252 recipe);
* the sampled solenoid field against an independent Biot-Savart sum of
  current loops (complete elliptic integrals from scipy).
"""

import ctypes as C
import math

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host

GRAM = 1.0 / 1.782662e-24
CM3 = 1e3
MUON_MASS = float(np.float32(np.float32(105.6583755) * np.float32(1e-3)))


def slab(X0, L0, A, Z, rho_g_cm3, thickness):
    s = capi.MaterialSlab()
    s.X0, s.L0, s.A, s.Z = X0, L0, A, Z
    s.mass_rho = rho_g_cm3 * GRAM / CM3
    s.molar_rho = s.mass_rho / (A * GRAM)
    s.mean_exc = 16e-9 * Z ** 0.9
    s.thickness = thickness
    s.thickness_in_X0 = thickness / X0
    return s


SILICON = dict(X0=93.7, L0=465.2, A=28.0855, Z=14.0, rho_g_cm3=2.329)
ALUMINIUM = dict(X0=88.97, L0=397.0, A=26.98, Z=13.0, rho_g_cm3=2.699)


@pytest.mark.parametrize("mat,p,expected", [
    (SILICON, 0.1003, 2.608), (SILICON, 1.101, 1.803), (SILICON, 10.11, 2.177),
    (SILICON, 100.1, 2.451), (ALUMINIUM, 1.101, 1.744), (ALUMINIUM, 10.11, 2.097),
    (ALUMINIUM, 100.1, 2.360),
])
def test_bethe_bloch_pdg(mat, p, expected):
    s = slab(**mat, thickness=10.0)
    dedx, _, _ = oracle_host.material_probe(s, -1.0 / p, 10.0, 0.0)
    # GeV/mm -> MeV cm^2 / g
    value = dedx / s.mass_rho / (1e-3 * 1e2 / GRAM)
    assert abs(expected - value) / value < 0.05


def test_landau_fwhm_kat():
    # 10 GeV muon energy through 1.7 mm Si: FWHM 0.13 MeV (RPP2018 fig. 33.7)
    s = slab(**SILICON, thickness=1.7)
    E = 10.0
    p = math.sqrt(E * E - MUON_MASS ** 2)
    qop = -1.0 / p
    _, sigma_qop, _ = oracle_host.material_probe(s, qop, 1.7, 0.0)
    beta2 = p * p / E / E
    q2b2 = 1.0 / beta2
    sigma_e = sigma_qop / (math.sqrt(q2b2) * qop * qop)
    fwhm = sigma_e * 2.0 * math.sqrt(2.0) * math.sqrt(math.log(2.0))
    assert abs(0.13 - fwhm * 1e3) / (fwhm * 1e3) < 0.1


@pytest.mark.parametrize("p,x", [(0.5, 0.01), (1.0, 0.016), (10.0, 0.05)])
def test_highland_pdg(p, x):
    s = slab(**SILICON, thickness=1.0)
    _, _, theta0 = oracle_host.material_probe(s, -1.0 / p, 1.0, x)
    E = math.sqrt(p * p + MUON_MASS ** 2)
    beta = p / E
    pdg = 13.6e-3 / (beta * p) * math.sqrt(x) * (1.0 + 0.038 * math.log(x / beta ** 2))
    assert theta0 == pytest.approx(pdg, rel=1e-6)


def helix_at_z(pos, dirn, qop, bvec, z_target):
    """Analytic helix in a constant field, dT/ds = qop (T x B); returns the
    position and direction where z(s) = z_target (Newton on s)."""
    b = np.linalg.norm(bvec)
    bh = bvec / b
    k = qop * b
    t_par = np.dot(dirn, bh) * bh
    t_perp = dirn - t_par
    bxt = np.cross(bh, t_perp)

    def r(s):
        return pos + t_par * s + t_perp * math.sin(k * s) / k - bxt * (1 - math.cos(k * s)) / k

    def t(s):
        return t_par + t_perp * math.cos(k * s) - bxt * math.sin(k * s)

    s = (z_target - pos[2]) / dirn[2]
    for _ in range(50):
        f = r(s)[2] - z_target
        s -= f / t(s)[2]
        if abs(f) < 1e-13:
            break
    return r(s), t(s)


def test_rk4_against_helix():
    # No step constraint: rk_stepper.hpp:474-511 clips the step after the
    # error estimate and advances with the stages of the unclipped step, so a
    # binding constraint adds an O(h^2 k) error per clipped step; without one
    # every step is a consistent RKN4 step.
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.0, 60.0)
    bvec = np.array([0.0, 1.0 * da.TESLA, 0.0])
    field = da.MagneticField.constant(*bvec)
    cfg = da.PropagationConfig(estimate_scattering_noise=False)
    fp = da.generate_tracks("beam", 256, 42)
    out = oracle_host.propagate(det, field, cfg, fp)
    assert np.all(out["n_surfaces"] == 10)
    for i in range(fp.shape[1]):
        pos0, dir0, qop = fp[0:3, i], fp[4:7, i], fp[7, i]
        r, t = helix_at_z(pos0, dir0, qop, bvec, 1000.0)
        b = out["bound_params"][:, i]
        # relative position error |dr| / s (the reference bounds it by 1e-3)
        s_len = out["path_length"][i]
        assert math.hypot(b[0] - r[0], b[1] - r[1]) / s_len < 1e-5
        d = np.array([math.cos(b[2]) * math.sin(b[3]), math.sin(b[2]) * math.sin(b[3]),
                      math.cos(b[3])])
        assert np.linalg.norm(d - t) < 1e-7
        assert b[4] == qop


def _cov_full(c21):
    full = np.zeros((6, 6))
    k = 0
    for r in range(6):
        for c in range(r, 6):
            full[r, c] = full[c, r] = c21[k]
            k += 1
    return full


def test_covariance_transport_vs_finite_differences():
    """Bound start on plane 1, no material: the covariance at plane 10 is
    J C0 J^T with J the bound-to-bound Jacobian of the propagation; J from
    central differences of the oracle's own propagation map."""
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.0, 60.0)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.2 * da.TESLA)
    cfg = da.PropagationConfig(estimate_scattering_noise=False)
    start = [0.0, 0.0, 0.05, 0.08, -1.0 / 2.0, 0.0]  # loc0, loc1, phi, theta, qop, t
    h = np.array([1e-4, 1e-4, 1e-6, 1e-6, 1e-6, 1e-4])
    cols = 1 + 2 * 6
    bp = np.tile(np.array(start)[:, None], (1, cols))
    for j in range(6):
        bp[j, 1 + 2 * j] += h[j]
        bp[j, 2 + 2 * j] -= h[j]
    c0 = np.zeros((21, cols))
    for k in (0, 6, 11, 15, 18, 20):  # identity
        c0[k] = 1.0
    out = oracle_host.propagate(det, field, cfg, None, initial_cov=c0, bound_params=bp,
                                start_surface=np.zeros(cols, dtype=np.uint32))
    assert np.all(out["status"] == 0)
    # the start surface is recorded first (the sequencer sees it at i = 0)
    assert np.all(out["surfaces"][0] == 0) and np.all(out["surfaces"][1] == 1)
    assert np.all(out["n_surfaces"] == 10)
    J = np.zeros((6, 6))
    for j in range(6):
        J[:, j] = (out["bound_params"][:, 1 + 2 * j] - out["bound_params"][:, 2 + 2 * j]) / (2 * h[j])
    cov = _cov_full(out["bound_cov"][:, 0])
    ref = J @ J.T
    d = np.sqrt(np.diag(ref))
    rel = np.abs(cov - ref) / np.outer(d, d)
    assert np.max(rel) < 1e-5, rel


def _loop_field(a, z0, rho, z):
    """Field of a unit-current loop of radius a at height z0 (Jackson
    5.37, complete elliptic integrals), mu0 I / (2 pi) = 1."""
    from scipy.special import ellipe, ellipk

    dz = z - z0
    alpha2 = a * a + rho * rho + dz * dz - 2 * a * rho
    beta2 = a * a + rho * rho + dz * dz + 2 * a * rho
    k2 = 1.0 - alpha2 / beta2
    K, E = ellipk(k2), ellipe(k2)
    beta = math.sqrt(beta2)
    bz = (1.0 / beta) * (E * (a * a - rho * rho - dz * dz) / alpha2 + K)
    if rho == 0.0:
        br = 0.0
    else:
        br = (dz / (rho * beta)) * (E * (a * a + rho * rho + dz * dz) / alpha2 - K)
    return br, bz


def test_solenoid_against_biot_savart():
    from scipy.integrate import quad

    R, L, B0 = 1200.0, 3000.0, 2.0 * da.TESLA
    field = da.MagneticField.solenoid()

    def bs(rho, z):
        br = quad(lambda zp: _loop_field(R, zp, rho, z)[0], -L, L, limit=400, points=[z])[0]
        bz = quad(lambda zp: _loop_field(R, zp, rho, z)[1], -L, L, limit=400, points=[z])[0]
        return br, bz

    _, bz_c = bs(0.0, 0.0)
    scale = B0 / bz_c
    # grid nodes: the oracle returns the stored fp32 samples there
    pts = np.array([[0.0, 0.0, 0.0], [500.0, 0.0, 1000.0], [0.0, 800.0, -2500.0],
                    [-600.0, 600.0, 2900.0], [1000.0, 0.0, 3500.0], [300.0, -400.0, 0.0]])
    got = oracle_host.field_at(field, pts)
    for p, b in zip(pts, got):
        rho = math.hypot(p[0], p[1])
        br, bz = bs(rho, p[2])
        br, bz = br * scale, bz * scale
        bx = br * p[0] / rho if rho > 0 else 0.0
        by = br * p[1] / rho if rho > 0 else 0.0
        ref = np.array([bx, by, bz])
        assert np.max(np.abs(b - ref)) < 2e-6 * B0 + 1e-6 * np.max(np.abs(ref)), (p, b, ref)


def test_solenoid_on_axis_closed_form():
    field = da.MagneticField.solenoid()
    R, L, B0 = 1200.0, 3000.0, 2.0 * da.TESLA
    zs = np.arange(-3500.0, 3501.0, 250.0)
    pts = np.stack([np.zeros_like(zs), np.zeros_like(zs), zs], axis=1)
    got = oracle_host.field_at(field, pts)[:, 2]

    def onaxis(z):
        return (z + L) / math.hypot(z + L, R) - (z - L) / math.hypot(z - L, R)

    ref = np.array([B0 * onaxis(z) / onaxis(0.0) for z in zs])
    assert np.max(np.abs(got - ref)) < 1e-6 * B0


def test_field_interpolation_and_clamping():
    field = da.MagneticField.solenoid()
    # midpoint between grid nodes: trilinear interpolation of the samples
    # stays within the interpolation error bound of the smooth field
    mid = np.array([[25.0, 25.0, 25.0], [1025.0, 25.0, 2975.0]])
    nodes = oracle_host.field_at(field, mid + np.array([[-25.0, -25.0, -25.0]]))
    got = oracle_host.field_at(field, mid)
    assert np.all(np.abs(got - nodes) < 0.05 * 2.0 * da.TESLA)
    # outside the grid the value is clamped to the border
    out = oracle_host.field_at(field, np.array([[0.0, 0.0, 9000.0], [0.0, 0.0, 3500.0]]))
    assert np.array_equal(out[0], out[1])
