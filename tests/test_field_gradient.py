"""use_field_gradient / use_eloss_gradient (stepping::config,
stepping_config.hpp:27-48) in the oracle, no GPU.

With use_field_gradient the Runge-Kutta Jacobian gains the field gradient
dB/dr (central differences of +-0.1 mm around the end of the step,
rk_stepper.hpp:878-998).  That term only enters the derivatives of the
stages with respect to position, so what the mathematics fixes is the
position columns of the transported Jacobian: they must equal central finite
differences of the propagated parameters in a field with a gradient, and
without the term they must not.  The direction and q/p columns keep the
reference's approximate formula (no gradient terms in dk/dt, dk/dqop), so
only an improvement is asserted for the full covariance.

use_eloss_gradient adds d(dqop/ds)/dqop of the VOLUME material to the
Jacobian (rk_stepper.hpp:958-998); the toy and telescope detectors carry
surface material only, so the flag changes nothing.
"""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

START = (5.0, -3.0, 0.05, 0.08, -0.5, 0.0)


def _cov_full(c21):
    n = c21.shape[1]
    full = np.zeros((6, 6, n))
    k = 0
    for r in range(6):
        for c in range(r, 6):
            full[r, c] = c21[k]
            full[c, r] = c21[k]
            k += 1
    return full


def gradient_field(kind):
    """Sampled field on a 50 mm grid around a telescope along z: a 1 T By,
    a 0.2 T Bz, plus a linear gradient of one component."""
    xs = np.arange(-600, 601, 50.0)
    zs = np.arange(-100, 1201, 50.0)
    X, Y, Z = np.meshgrid(xs, xs, zs, indexing="ij")
    v = np.zeros(X.shape + (3,))
    v[..., 1] = 1.0 * da.TESLA
    v[..., 2] = 0.2 * da.TESLA
    if kind == "Bx(x)":
        v[..., 0] = 0.3 * da.TESLA * X / 500.0
    elif kind == "Bz(y)":
        v[..., 2] += 0.3 * da.TESLA * Y / 500.0
    return da.MagneticField.grid(v, (-600, -600, -100), (50, 50, 50), name=kind)


@pytest.fixture(scope="module")
def telescope_z():
    return da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.0, 60.0)


def jacobian_check(det, field, grad, cols):
    """Transported covariance of the unit start covariance restricted to
    `cols` (i.e. J[:, cols] J[:, cols]^T) against central finite differences
    of the propagated bound parameters; error relative to the largest
    element."""
    cfg = da.PropagationConfig(estimate_scattering_noise=False, use_field_gradient=grad)
    h = np.array([1e-3, 1e-3, 1e-6, 1e-6, 1e-6, 1e-4])
    nc = 1 + 2 * len(cols)
    bp = np.tile(np.array(START)[:, None], (1, nc))
    for k, j in enumerate(cols):
        bp[j, 1 + 2 * k] += h[j]
        bp[j, 2 + 2 * k] -= h[j]
    c0 = np.zeros((21, nc))
    diag = (0, 6, 11, 15, 18, 20)
    for j in cols:
        c0[diag[j]] = 1.0
    out = oracle_host.propagate(det, field, cfg, None, initial_cov=c0, bound_params=bp,
                                start_surface=np.zeros(nc, dtype=np.uint32))
    assert (out["status"] == 0).all() and (out["n_surfaces"] == out["n_surfaces"][0]).all()
    J = np.zeros((6, len(cols)))
    for k, j in enumerate(cols):
        J[:, k] = (out["bound_params"][:, 1 + 2 * k] - out["bound_params"][:, 2 + 2 * k]) / (2 * h[j])
    cov = _cov_full(out["bound_cov"][:, :1])[:, :, 0]
    ref = J @ J.T
    return np.abs(cov - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("kind", ["Bx(x)", "Bz(y)"])
def test_position_columns_match_finite_differences(telescope_z, kind):
    field = gradient_field(kind)
    e_grad = jacobian_check(telescope_z, field, True, (0, 1))
    e_plain = jacobian_check(telescope_z, field, False, (0, 1))
    print(f"{kind}: position columns vs FD: with gradient {e_grad:.1e}, without {e_plain:.1e}")
    assert e_grad < 1e-5
    assert e_plain > 1e-3


@pytest.mark.parametrize("kind", ["Bx(x)", "Bz(y)"])
def test_full_jacobian_improves(telescope_z, kind):
    field = gradient_field(kind)
    cols = (0, 1, 2, 3, 4)
    assert jacobian_check(telescope_z, field, True, cols) < jacobian_check(telescope_z, field, False, cols)


def test_no_effect_in_a_constant_field(telescope_z):
    """A sampled constant field has a zero gradient: identical results."""
    field = gradient_field("const")
    fp = da.generate_tracks("beam", 64, 42)
    cfg = dict(step_constraint=50.0, estimate_scattering_noise=False)
    a = oracle_host.propagate(telescope_z, field, da.PropagationConfig(**cfg), fp)
    b = oracle_host.propagate(telescope_z, field, da.PropagationConfig(use_field_gradient=True, **cfg), fp)
    for k in ("bound_params", "free_params", "n_steps", "status", "n_surfaces"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_allclose(b["bound_cov"], a["bound_cov"], rtol=1e-12, atol=1e-300)


def test_eloss_gradient_without_volume_material(toy7, solenoid):
    fp = da.generate_tracks("toy", 256, 43)
    cfg = dict(search_window=(3, 3))
    a = oracle_host.propagate(toy7, solenoid, da.PropagationConfig(**cfg), fp)
    b = oracle_host.propagate(toy7, solenoid, da.PropagationConfig(use_eloss_gradient=True, **cfg), fp)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
