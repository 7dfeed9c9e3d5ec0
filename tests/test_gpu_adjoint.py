"""Adjoint kernel (dtr_adjoint_forward / dtr_adjoint_backward): gradient of
the chi2 of hit residuals with respect to the bound start parameters.

Workload: tracks on the toy detector's beam pipe in the solenoid map (pT
log-uniform 0.5-50 GeV, |eta| < 2.5), hits = true intersections of the
1-sigma perturbed start parameters (from the oracle) plus Gaussian smear
(10 um, 50 um).

Gradient comparisons use the error in units of the start uncertainties:
    e = max_i |g_i - r_i| s_i / max_i |r_i| s_i
(s = 10 um, 10 um, 1 mrad, 1 mrad, 1 % of q/p), i.e. the error of the
predicted chi2 change for 1-sigma moves, relative to the largest.

Contracts:
* exact adjoint vs central finite differences (h = 1e-6 relative) of the
  ORACLE's FP64 chi2: e <= 1e-3 for FP64 and mixed precision.  A track whose
  finite difference is itself unstable (h and h/2 disagree by more than
  1e-4: a step-control or on-surface decision flips inside +-h) is excluded;
  at most 3 % of tracks may be.
* transport adjoint vs the gradient of the chain of the reference's
  transport Jacobians (written by the forward pass): e <= 1e-6 (FP64),
  1e-4 (mixed: FP32 Jacobians).
* chi2 and the sensitive-surface parameters of the forward pass equal the
  oracle's (FP64).
* the result does not depend on the checkpoint interval K (bit-identical
  replay).
"""

import numpy as np
import pytest

import detray_amd as da
from detray_amd import adjoint as adj
from tests import oracle_host

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

MAX_HITS = 32
CONFIG = da.PropagationConfig(search_window=(3, 3))


def scaled_error(g, ref, sig):
    d = (np.abs(g[:5] - ref[:5]) * sig[:5]).max(axis=0)
    m = (np.abs(ref[:5]) * sig[:5]).max(axis=0)
    return d / np.maximum(m, 1e-300)


@pytest.fixture(scope="module")
def workload(toy7, solenoid):
    n = 1024
    b0, start = adj.beampipe_starts(n, seed=46)
    truth = adj.perturbed(b0, seed=47)
    rec = oracle_host.propagate_hits(toy7, solenoid, CONFIG, truth, start, MAX_HITS)
    rng = np.random.default_rng(48)
    meas_loc = rec["hit_loc"] + rng.normal(size=rec["hit_loc"].shape) * np.array(
        [adj.SIGMA_LOC0, adj.SIGMA_LOC1])[:, None, None]
    meas_sf = rec["hit_surface"].copy()
    meas_n = rec["n_hits"].copy()
    hits = da.Hits(torch.from_numpy(meas_sf.view(np.int32)).cuda(), torch.from_numpy(meas_loc).cuda(),
                   torch.from_numpy(meas_n.view(np.int32)).cuda())
    return dict(n=n, b0=b0, start=start, sig=adj.start_sigma(b0), hits=hits, meas_sf=meas_sf,
                meas_loc=meas_loc, meas_n=meas_n, det=toy7, field=solenoid)


@pytest.fixture(scope="module")
def oracle_fd(workload):
    w = workload
    args = (w["det"], w["field"], CONFIG, w["b0"], w["start"], w["meas_sf"], w["meas_loc"], w["meas_n"],
            adj.SIGMA_LOC0, adj.SIGMA_LOC1)
    fd = oracle_host.chi2_gradient_fd(*args, rel=1e-6)
    fd_half = oracle_host.chi2_gradient_fd(*args, rel=5e-7)
    stable = scaled_error(fd_half, fd, w["sig"]) <= 1e-4
    return fd, stable


def run(w, precision, model="exact", every=32, max_steps=512, max_events=64, transport=False, sort="eta_phi",
        waves=2, block=128):
    prop = da.Chi2Propagator(CONFIG, da.AdjointConfig(max_steps=max_steps, max_events=max_events,
                                                      checkpoint_every=every, precision=precision, model=model,
                                                      sort=sort, waves_per_simd=waves, block_size=block))
    bound = torch.from_numpy(w["b0"]).cuda()
    start = torch.from_numpy(w["start"]).cuda()
    fwd = prop.forward(w["det"], w["field"], bound, start, w["hits"], transport_gradient=transport)
    g = prop.backward(w["det"], w["field"], bound, start, w["hits"], fwd)
    torch.cuda.synchronize()
    return dict(chi2=fwd.chi2.cpu().numpy(), status=fwd.status.cpu().numpy(), grad=g.cpu().numpy(),
                gt=fwd.grad_transport.cpu().numpy() if transport else None)


def test_forward_matches_oracle(workload):
    w = workload
    r = run(w, "fp64")
    assert (r["status"] == 0).all()
    ref = oracle_host.chi2(oracle_host.propagate_hits(w["det"], w["field"], CONFIG, w["b0"], w["start"], MAX_HITS),
                           w["meas_sf"], w["meas_loc"], w["meas_n"], adj.SIGMA_LOC0, adj.SIGMA_LOC1)
    np.testing.assert_allclose(r["chi2"], ref, rtol=1e-9, atol=1e-9)


def test_recorded_hits_match_oracle(workload):
    w = workload
    prop = da.Chi2Propagator(CONFIG, da.AdjointConfig(precision="fp64"))
    truth = adj.perturbed(w["b0"], seed=47)
    rec = prop.record_hits(w["det"], w["field"], torch.from_numpy(truth).cuda(),
                           torch.from_numpy(w["start"]).cuda(), MAX_HITS)
    torch.cuda.synchronize()
    ref = oracle_host.propagate_hits(w["det"], w["field"], CONFIG, truth, w["start"], MAX_HITS)
    n_hits = rec.n_hits.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(n_hits, ref["n_hits"])
    sf = rec.surface.cpu().numpy().view(np.uint32)
    loc = rec.loc.cpu().numpy()
    for i in range(w["n"]):
        k = n_hits[i]
        np.testing.assert_array_equal(sf[:k, i], ref["hit_surface"][:k, i])
        np.testing.assert_allclose(loc[:, :k, i], ref["hit_loc"][:, :k, i], rtol=0, atol=1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_exact_adjoint_matches_oracle_finite_differences(workload, oracle_fd, precision):
    w = workload
    fd, stable = oracle_fd
    r = run(w, precision)
    ok = r["status"] == 0
    assert ok.mean() > 0.99
    assert stable.mean() >= 0.97
    e = scaled_error(r["grad"], fd, w["sig"])[ok & stable]
    print(f"[{precision}] exact adjoint vs oracle FD: max {e.max():.2e} median {np.median(e):.2e} "
          f"(unstable FD tracks excluded: {(~stable).sum()})")
    assert e.max() <= 1e-3


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-4)])
def test_transport_adjoint_matches_jacobian_chain(workload, precision, tol):
    w = workload
    r = run(w, precision, model="transport", transport=True)
    ok = r["status"] == 0
    e = scaled_error(r["grad"], r["gt"], w["sig"])[ok]
    print(f"[{precision}] transport adjoint vs Jacobian chain: max {e.max():.2e} median {np.median(e):.2e}")
    assert e.max() <= tol


def test_jacobian_chain_lacks_field_and_eloss_gradient(workload, oracle_fd):
    """The reference's Jacobians leave out the field gradient and the
    energy-loss derivative: against finite differences their gradient is
    visibly worse than the exact adjoint's."""
    w = workload
    fd, stable = oracle_fd
    r = run(w, "fp64", model="transport", transport=True)
    ok = (r["status"] == 0) & stable
    e_chain = scaled_error(r["gt"], fd, w["sig"])[ok]
    r2 = run(w, "fp64")
    e_exact = scaled_error(r2["grad"], fd, w["sig"])[ok]
    assert np.median(e_chain) > 2.0 * np.median(e_exact)


def test_checkpoint_interval_does_not_change_the_gradient(workload):
    w = workload
    ref = run(w, "mixed", every=32)
    for every in (1, 7, 512):
        r = run(w, "mixed", every=every)
        np.testing.assert_array_equal(r["status"], ref["status"])
        np.testing.assert_array_equal(r["grad"], ref["grad"])


def test_launch_and_order_do_not_change_the_gradient(workload):
    """Track order (tape columns follow the order), register budget and
    block size only change the schedule: bit-identical results."""
    w = workload
    ref = run(w, "mixed")
    for kw in (dict(sort="none"), dict(waves=1), dict(block=64), dict(block=256, sort="none")):
        r = run(w, "mixed", **kw)
        np.testing.assert_array_equal(r["chi2"], ref["chi2"])
        np.testing.assert_array_equal(r["grad"], ref["grad"])


def test_covariance_is_skipped_without_noise_estimation(workload):
    """Without noise estimation the forward pass does not carry the
    covariance (unless the Jacobian chain is asked for); the trajectory, chi2
    and gradient are the same either way."""
    w = workload
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    prop = da.Chi2Propagator(cfg, da.AdjointConfig(precision="mixed"))
    bound = torch.from_numpy(w["b0"]).cuda()
    start = torch.from_numpy(w["start"]).cuda()
    f0 = prop.forward(w["det"], w["field"], bound, start, w["hits"])
    g0 = prop.backward(w["det"], w["field"], bound, start, w["hits"], f0)
    f1 = prop.forward(w["det"], w["field"], bound, start, w["hits"], transport_gradient=True)
    g1 = prop.backward(w["det"], w["field"], bound, start, w["hits"], f1)
    torch.cuda.synchronize()
    assert torch.equal(f0.chi2, f1.chi2)
    assert torch.equal(g0, g1)


def test_tape_overflow_is_reported(workload):
    w = workload
    r = run(w, "mixed", every=4, max_steps=8)
    assert (r["status"] == 1).mean() > 0.9
    assert np.isnan(r["grad"][:, r["status"] != 0]).all()
    r = run(w, "mixed", max_events=2)
    assert (r["status"] == 2).mean() > 0.9
    assert np.isnan(r["grad"][:, r["status"] != 0]).all()


def test_autograd_function(workload):
    w = workload
    bound = torch.from_numpy(w["b0"]).cuda().requires_grad_(True)
    start = torch.from_numpy(w["start"]).cuda()
    prop = da.Chi2Propagator(CONFIG, da.AdjointConfig(precision="mixed"))
    c = prop.chi2(w["det"], w["field"], bound, start, w["hits"])
    weights = torch.linspace(0.5, 1.5, w["n"], dtype=torch.float64, device="cuda")
    (g,) = torch.autograd.grad((weights * c).sum(), bound)
    ref = run(w, "mixed")
    np.testing.assert_allclose(c.detach().cpu().numpy(), ref["chi2"], rtol=0, atol=0)
    np.testing.assert_allclose(g.cpu().numpy(), ref["grad"] * weights.cpu().numpy()[None, :], rtol=1e-15, atol=0)


def test_empty_batch_and_errors(workload):
    w = workload
    prop = da.Chi2Propagator(CONFIG, da.AdjointConfig())
    empty = torch.zeros((6, 0), dtype=torch.float64, device="cuda")
    es = torch.zeros((0,), dtype=torch.int32, device="cuda")
    fwd = prop.forward(w["det"], w["field"], empty, es)
    assert fwd.chi2.numel() == 0
    with pytest.raises(ValueError):
        prop.forward(w["det"], w["field"], torch.zeros((5, 4), dtype=torch.float64, device="cuda"),
                     torch.zeros(4, dtype=torch.int32, device="cuda"))
