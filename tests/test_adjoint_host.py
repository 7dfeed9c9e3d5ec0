"""Host side of the differentiation form (no GPU): the adjoint's C ABI
(defaults, tape size, argument checks), the workload generators, and the
oracle's chi2 / finite-difference machinery that the GPU adjoint is checked
against, pinned on straight lines through a telescope where chi2 and its
gradient have a closed form."""

import ctypes as C
import math

import numpy as np
import pytest

import detray_amd as da
from detray_amd import adjoint as adj
from detray_amd import capi
from tests import oracle_host


def test_default_adjoint_config():
    a = capi.AdjointConfig()
    capi.load().dtr_default_adjoint(C.byref(a))
    assert (a.max_steps, a.max_events, a.checkpoint_every) == (512, 64, 32)
    assert a.precision == capi.PRECISION_MIXED and a.model == capi.ADJOINT_EXACT
    assert a.block_size == 64 and a.waves_per_simd == 2
    assert a.sigma_loc0 == 0.010 and a.sigma_loc1 == 0.050
    c = da.AdjointConfig().to_c()
    for f, _ in capi.AdjointConfig._fields_:
        assert getattr(c, f) == getattr(a, f), f
    assert da.AdjointConfig(precision="fp64", model="transport").to_c().model == capi.ADJOINT_TRANSPORT


@pytest.mark.parametrize("ms,me,k", [(512, 64, 32), (100, 7, 9), (8, 1, 8)])
def test_tape_size(ms, me, k):
    a = da.AdjointConfig(max_steps=ms, max_events=me, checkpoint_every=k).to_c()
    n = 1000
    doubles = 2 * ms + 7 * (ms // k + 1) + (10 + 6 + 7) * me + 7 * k + 36
    words = 4 * me + ms + 3  # per event: step, surface, hit, flags; per step: its target surface
    assert capi.load().dtr_adjoint_tape_size(C.byref(a), n) == n * (8 * doubles + 4 * words)
    bad = da.AdjointConfig(checkpoint_every=0).to_c()
    assert capi.load().dtr_adjoint_tape_size(C.byref(bad), n) == 0


def test_adjoint_entry_points_reject_bad_arguments():
    lib = capi.load()
    cfg = da.PropagationConfig().to_c()
    a = da.AdjointConfig().to_c()
    b = capi.AdjointBatch(4, 0, None, None, None, 0)
    assert lib.dtr_adjoint_forward(None, None, C.byref(cfg), C.byref(a), C.byref(b), None, None, None, None, None,
                                   None) == 1
    assert lib.dtr_adjoint_backward(None, None, C.byref(cfg), C.byref(a), C.byref(b), None, None,
                                    None) == 1


def test_beampipe_starts():
    b, start = adj.beampipe_starts(20000, seed=46)
    assert b.shape == (6, 20000) and (start == adj.BEAMPIPE_SURFACE).all()
    eta = -np.log(np.tan(b[3] / 2))
    assert np.abs(eta).max() <= 2.5 + 1e-12
    pt = np.sin(b[3]) / np.abs(b[4])
    assert pt.min() >= 0.5 * (1 - 1e-12) and pt.max() <= 50.0 * (1 + 1e-12)
    # outgoing: momentum azimuth within a few 10 mrad of the position azimuth
    dphi = (b[2] - b[0] + math.pi) % (2 * math.pi) - math.pi
    assert np.abs(dphi).max() < 0.08
    assert set(np.unique(np.sign(b[4]))) == {-1.0, 1.0}
    # the truth perturbation is 1 sigma on average
    t = adj.perturbed(b, seed=47)
    s = adj.start_sigma(b)
    z = (t - b)[:5] / s[:5]
    assert np.allclose(z.std(axis=1), 1.0, atol=0.03)


def _straight_line_telescope():
    det = da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.0, 60.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    rng = np.random.default_rng(7)
    n = 64
    b = np.zeros((6, n))
    b[0] = rng.uniform(-20, 20, n)
    b[1] = rng.uniform(-20, 20, n)
    b[2] = rng.uniform(-math.pi, math.pi, n)
    b[3] = rng.uniform(0.01, 0.2, n)
    b[4] = rng.choice([-1.0, 1.0], n) / rng.uniform(1.0, 10.0, n)
    return det, field, b, np.zeros(n, dtype=np.uint32)


def _closed_form_locs(b, z):
    dz = z[:, None] - 100.0
    t = np.tan(b[3])
    return np.stack([b[0] + dz * t * np.cos(b[2]), b[1] + dz * t * np.sin(b[2])])  # [2, planes, n]


def test_oracle_hits_on_straight_lines():
    det, field, b, start = _straight_line_telescope()
    rec = oracle_host.propagate_hits(det, field, da.PropagationConfig(), b, start, 16)
    assert (rec["n_hits"] == 10).all()
    np.testing.assert_array_equal(rec["hit_surface"][:10], np.arange(10, dtype=np.uint32)[:, None].repeat(b.shape[1], 1))
    z = 100.0 * np.arange(1, 11)
    np.testing.assert_allclose(rec["hit_loc"][:, :10], _closed_form_locs(b, z), rtol=0, atol=1e-9)


def test_oracle_chi2_and_fd_gradient_closed_form():
    """chi2 of straight lines against measurements and its gradient in
    closed form; the oracle's chi2 and central finite differences must
    reproduce both (this pins the machinery the GPU adjoint is checked
    against)."""
    det, field, b, start = _straight_line_telescope()
    n = b.shape[1]
    z = 100.0 * np.arange(1, 11)
    rng = np.random.default_rng(8)
    meas = _closed_form_locs(b, z) + rng.normal(size=(2, 10, n)) * np.array([0.01, 0.05])[:, None, None]
    meas_loc = np.zeros((2, 16, n))
    meas_loc[:, :10] = meas
    # drop the 4th hit of every other track: the matching skips it
    meas_sf = np.full((16, n), 0xFFFFFFFF, dtype=np.uint32)
    meas_sf[:10] = np.arange(10, dtype=np.uint32)[:, None]
    meas_n = np.full(n, 10, dtype=np.uint32)
    meas_sf[3, ::2] = 0xFFFFFFFE
    s0, s1 = 0.01, 0.05
    r = _closed_form_locs(b, z) - meas
    use = np.ones((10, n), dtype=bool)
    use[3, ::2] = False
    chi2 = ((r[0] ** 2 / s0 ** 2 + r[1] ** 2 / s1 ** 2) * use).sum(axis=0)
    rec = oracle_host.propagate_hits(det, field, da.PropagationConfig(), b, start, 16)
    np.testing.assert_allclose(oracle_host.chi2(rec, meas_sf, meas_loc, meas_n, s0, s1), chi2, rtol=1e-9)
    # d loc_k / d(loc0, loc1, phi, theta, q/p)
    dz = z[:, None] - 100.0
    t, sec2 = np.tan(b[3]), 1.0 / np.cos(b[3]) ** 2
    cp, sp = np.cos(b[2]), np.sin(b[2])
    w0 = 2 * r[0] / s0 ** 2 * use
    w1 = 2 * r[1] / s1 ** 2 * use
    g = np.zeros((6, n))
    g[0] = w0.sum(axis=0)
    g[1] = w1.sum(axis=0)
    g[2] = (w0 * dz * t * -sp + w1 * dz * t * cp).sum(axis=0)
    g[3] = (w0 * dz * sec2 * cp + w1 * dz * sec2 * sp).sum(axis=0)
    fd = oracle_host.chi2_gradient_fd(det, field, da.PropagationConfig(), b, start, meas_sf, meas_loc, meas_n,
                                      s0, s1, rel=1e-6, max_hits=16)
    scale = np.abs(g[:4]).max(axis=0)
    assert (np.abs(fd[:4] - g[:4]).max(axis=0) / scale).max() < 1e-6
    assert np.abs(fd[4]).max() == 0.0  # no field: q/p does not enter
