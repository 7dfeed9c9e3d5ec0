"""Device path (libdetray_amd.so, HIP) against the oracle on the same seeded
inputs: identical surface sequences, statuses and step/trial counts; bound
parameters and covariance within stated tolerances.

Tolerances (FP64 vs FP64, same algorithm, different summation order in the
Jacobian products and a different libm for sin/cos/atan2/log):
  * bound parameters: |d| <= 1e-9 * max(|ref|, 1) per element
  * covariance: |dC_ij| <= 1e-7 * sqrt(C_ii C_jj)
"""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

BOUND_TOL = 1e-9
COV_TOL = 1e-7


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


def run_both(det, field, cfg, fp, initial_cov=None):
    prop = da.Propagator(cfg)
    dev_cov = torch.from_numpy(initial_cov).cuda() if initial_cov is not None else None
    res = prop.propagate(det, field, torch.from_numpy(fp).cuda(), initial_cov=dev_cov)
    torch.cuda.synchronize()
    got = {
        "bound_params": res.bound_params.cpu().numpy(),
        "bound_cov": res.bound_cov.cpu().numpy(),
        "free_params": res.free_params.cpu().numpy(),
        "path_length": res.path_length.cpu().numpy(),
        "surfaces": res.surfaces.cpu().numpy().view(np.uint32),
        "n_surfaces": res.n_surfaces.cpu().numpy().view(np.uint32),
        "n_steps": res.n_steps.cpu().numpy().view(np.uint32),
        "n_trials": res.n_trials.cpu().numpy().view(np.uint32),
        "status": res.status.cpu().numpy(),
    }
    ref = oracle_host.propagate(det, field, cfg, fp, initial_cov=initial_cov)
    return got, ref


def assert_parity(got, ref, check_cov=True):
    n = ref["status"].shape[0]
    assert np.array_equal(got["status"], ref["status"])
    assert np.array_equal(got["n_surfaces"], ref["n_surfaces"])
    ms = got["surfaces"].shape[0]
    for i in range(n):
        k = min(int(ref["n_surfaces"][i]), ms)
        assert np.array_equal(got["surfaces"][:k, i], ref["surfaces"][:k, i]), f"track {i}"
    # Tracks that leave the world are compared step for step.  Tracks stopped
    # by the path limit or by energy loss can split their last steps
    # differently (the step-size control amplifies last-bit differences of
    # sin/cos/atan2 between libm and the device math library over thousands
    # of steps); for them the surface sequence and the bound parameters at
    # the last surface are the contract.
    ex = ref["status"] == 0
    assert np.array_equal(got["n_steps"][ex], ref["n_steps"][ex])
    assert np.array_equal(got["n_trials"][ex], ref["n_trials"][ex])
    scale = np.maximum(np.abs(ref["bound_params"]), 1.0)
    assert np.max(np.abs(got["bound_params"] - ref["bound_params"]) / scale, initial=0.0) <= BOUND_TOL
    fscale = np.maximum(np.abs(ref["free_params"][:, ex]), 1.0)
    assert np.max(np.abs(got["free_params"][:, ex] - ref["free_params"][:, ex]) / fscale,
                  initial=0.0) <= BOUND_TOL
    if check_cov:
        cg, cr = _cov_full(got["bound_cov"]), _cov_full(ref["bound_cov"])
        d = np.sqrt(np.abs(np.einsum("iin->in", cr)))
        norm = np.einsum("in,jn->ijn", d, d) + 1e-300
        has = ref["n_surfaces"] > 0
        rel = np.abs(cg - cr)[:, :, has] / norm[:, :, has]
        assert rel.size == 0 or np.max(rel) <= COV_TOL, np.max(rel)


def test_telescope_const_field(telescope):
    cfg = da.PropagationConfig(step_constraint=50.0, estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0)
    fp = da.generate_tracks("beam", 4096, 42)
    got, ref = run_both(telescope, field, cfg, fp)
    assert np.all(ref["n_surfaces"] == 10)
    assert_parity(got, ref)


def test_telescope_initial_covariance(telescope):
    cfg = da.PropagationConfig(step_constraint=50.0, estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0)
    fp = da.generate_tracks("beam", 1024, 42)
    n = fp.shape[1]
    c0 = np.zeros((21, n))
    diag = {0: 0.01 ** 2, 6: 0.01 ** 2, 11: 1e-6, 15: 1e-6, 18: (0.01 * fp[7]) ** 2, 20: 1.0}
    for k, v in diag.items():
        c0[k] = v
    got, ref = run_both(telescope, field, cfg, fp, initial_cov=c0)
    assert_parity(got, ref)


@pytest.mark.parametrize("noise", [False, True])
def test_toy_const_field(toy7, noise):
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=noise)
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    fp = da.generate_tracks("toy", 8192, 43)
    got, ref = run_both(toy7, field, cfg, fp)
    assert_parity(got, ref)


def test_toy_solenoid(toy7, solenoid):
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = da.generate_tracks("toy", 8192, 44)
    got, ref = run_both(toy7, solenoid, cfg, fp)
    assert_parity(got, ref)


def test_toy_loopers(toy7, solenoid):
    # the batching-form spectrum: pT log-uniform 0.1-100 GeV, |eta| < 3
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = da.generate_tracks("toy", 4096, 45, pt_min=0.1, pt_max=100.0, eta_min=-3.0, eta_max=3.0)
    got, ref = run_both(toy7, solenoid, cfg, fp)
    assert_parity(got, ref)


def test_toy_path_limit(toy7, solenoid):
    # curling tracks (2R < beampipe volume) reach the 5 m path limit
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    fp = da.generate_tracks("toy", 512, 48, pt_min=0.02, pt_max=0.04, eta_min=-0.3, eta_max=0.3)
    got, ref = run_both(toy7, solenoid, cfg, fp)
    assert np.any(ref["status"] == 2)
    assert_parity(got, ref)


def test_sequencer_overflow(toy7):
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False, max_surfaces=3)
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    fp = da.generate_tracks("toy", 1024, 46)
    got, ref = run_both(toy7, field, cfg, fp)
    assert np.any(ref["status"] == 3)
    assert_parity(got, ref)


def test_no_material_chain(toy7):
    cfg = da.PropagationConfig(search_window=(3, 3), use_material=False,
                               do_covariance_transport=False, estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    fp = da.generate_tracks("toy", 2048, 47)
    got, ref = run_both(toy7, field, cfg, fp)
    assert_parity(got, ref, check_cov=False)


def test_empty_batch(toy7):
    prop = da.Propagator(da.PropagationConfig())
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    out = prop.propagate(toy7, field, torch.zeros((8, 0), dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    assert out.status.numel() == 0


def test_kat_navigation_ray(toy3):
    # caching_navigator.cpp:128-292: ray (1,1,0) through the default toy
    # detector, straight line (B = 0) -> surface sequence of the KAT
    cfg = da.PropagationConfig(search_window=(3, 3), use_material=False,
                               estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    fp = np.zeros((8, 1))
    fp[4, 0] = fp[5, 0] = 1.0 / np.sqrt(2.0)
    fp[7, 0] = -1.0 / np.sqrt(2.0)
    got, ref = run_both(toy3, field, cfg, fp)
    seq = got["surfaces"][: got["n_surfaces"][0], 0].tolist()
    assert seq == [0, 475, 459, 476, 460, 829, 797, 830, 798, 1438, 1386, 2372, 2294]


def test_bound_start_parity(telescope):
    # bound start on plane 1 (propagator::state(bound_track_parameters, ...))
    cfg = da.PropagationConfig(estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.2 * da.TESLA)
    n = 2048
    rng = np.random.default_rng(7)
    bp = np.zeros((6, n))
    bp[0] = rng.uniform(-50, 50, n)
    bp[1] = rng.uniform(-50, 50, n)
    bp[2] = rng.uniform(-np.pi, np.pi, n)
    bp[3] = rng.uniform(0.0, 0.2, n)
    bp[4] = rng.choice([-1.0, 1.0], n) / rng.uniform(1.0, 10.0, n)
    c0 = np.zeros((21, n))
    for k in (0, 6, 11, 15, 18, 20):
        c0[k] = 1e-4
    ss = np.zeros(n, dtype=np.int32)
    prop = da.Propagator(cfg)
    res = prop.propagate(telescope, field, bound_params=torch.from_numpy(bp).cuda(),
                         start_surface=torch.from_numpy(ss).cuda(),
                         initial_cov=torch.from_numpy(c0).cuda())
    torch.cuda.synchronize()
    got = {k: getattr(res, k).cpu().numpy() for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")}
    for k in ("surfaces", "n_surfaces", "n_steps", "n_trials"):
        got[k] = got[k].view(np.uint32)
    ref = oracle_host.propagate(telescope, field, cfg, None, initial_cov=c0, bound_params=bp,
                                start_surface=ss.astype(np.uint32))
    assert_parity(got, ref)
