"""Batched kernels (dtr_propagate_ex) against the oracle on the same seeded
inputs.

FP64 precision: the same contract as the plain kernel (identical surface
sequences, statuses and step counts; bound parameters within 1e-9, covariance
within 1e-7 of sqrt(C_ii C_jj)).

Mixed precision (FP32 stages, field, transport and material noise terms;
FP64 track state, path, error estimate, mean energy loss and navigation):
identical surface sequences and statuses; bound parameters within 1e-4 in
the reference's own host/device convention (propagator_test.hpp:70,
204-213: position differences relative to the path length), i.e.
  |d loc| <= 1e-4 * max(path, 1 mm), |d angle| <= 1e-4 * max(|angle|, 1),
  |d q/p| <= 1e-4 * |q/p|,
and covariance within 1e-4 of sqrt(C_ii C_jj) (2e-2 for tracks brought to
rest in material; 1e-3 for tracks ending within 0.02 rad of the z axis, where
the (phi, theta) parameters are singular: the phi variance grows as
1/sin^2(theta) and its FP32 rounding reaches the other elements; such tracks
only occur in the telescopes along z, at sin(theta) >= 0.02 the error there
stays below 1e-5).  (FP32 stage rounding can
flip an accept/reject decision of the adaptive step control, so a track may
take one step more or less than in FP64 and land a truncation error,
~1e-4 mm, away: loc differences are therefore not held to 1e-4 mm.)

Every schedule (one work-item per track, or persistent waves over a work
queue, with any permutation of the track order) must give the same result
for a track.
"""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host
from tests.test_gpu_parity import _cov_full, assert_parity

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

MIXED_BOUND_TOL = 1e-4
MIXED_COV_TOL = 1e-4
# tracks brought to rest in material (status abort after ~1000 steps of a
# 20-40 MeV muon): their last covariance is dominated by the straggling and
# scattering terms of the surface where p -> 0, which change by percents
# under last-bit perturbations of the trajectory
MIXED_COV_TOL_STOPPED = 2e-2
# tracks ending within 0.02 rad of the z axis (see the module docstring)
MIXED_COV_TOL_AXIAL = 1e-3
AXIAL_SIN_THETA = 0.02


def bound_error_scale(ref):
    """Per-element scale of the mixed-precision bound-parameter error."""
    path = np.maximum(np.abs(ref["path_length"]), 1.0)
    bp = np.abs(ref["bound_params"])
    scale = np.empty_like(bp)
    scale[0] = path
    scale[1] = path
    scale[2] = np.maximum(bp[2], 1.0)
    scale[3] = np.maximum(bp[3], 1.0)
    scale[4] = np.where(bp[4] > 0, bp[4], 1.0)
    scale[5] = np.maximum(bp[5], 1.0)
    return scale


def run_launch(det, field, cfg, fp, launch, initial_cov=None, bound=None, start=None):
    prop = da.Propagator(cfg, launch)
    kw = {}
    if bound is not None:
        kw = dict(bound_params=torch.from_numpy(bound).cuda(),
                  start_surface=torch.from_numpy(start.astype(np.int32)).cuda())
        tracks = None
    else:
        tracks = torch.from_numpy(fp).cuda()
    dev_cov = torch.from_numpy(initial_cov).cuda() if initial_cov is not None else None
    res = prop.propagate(det, field, tracks, initial_cov=dev_cov, **kw)
    torch.cuda.synchronize()
    return {
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


def assert_mixed_parity(got, ref):
    n = ref["status"].shape[0]
    assert np.array_equal(got["status"], ref["status"])
    assert np.array_equal(got["n_surfaces"], ref["n_surfaces"])
    ms = got["surfaces"].shape[0]
    for i in range(n):
        k = min(int(ref["n_surfaces"][i]), ms)
        assert np.array_equal(got["surfaces"][:k, i], ref["surfaces"][:k, i]), f"track {i}"
    err = np.abs(got["bound_params"] - ref["bound_params"]) / bound_error_scale(ref)
    assert np.max(err, initial=0.0) <= MIXED_BOUND_TOL, np.max(err, axis=1)
    cg, cr = _cov_full(got["bound_cov"]), _cov_full(ref["bound_cov"])
    d = np.sqrt(np.abs(np.einsum("iin->in", cr)))
    norm = np.einsum("in,jn->ijn", d, d) + 1e-300
    stopped = ref["status"] == 1
    axial = ~stopped & (np.abs(np.sin(ref["bound_params"][3])) < AXIAL_SIN_THETA)
    for sel, tol in ((~stopped & ~axial, MIXED_COV_TOL), (stopped, MIXED_COV_TOL_STOPPED),
                     (axial, MIXED_COV_TOL_AXIAL)):
        has = (ref["n_surfaces"] > 0) & sel
        rel = np.abs(cg - cr)[:, :, has] / norm[:, :, has]
        assert rel.size == 0 or np.max(rel) <= tol, (np.max(rel), tol)


def check(got, ref, precision):
    if precision == "fp64":
        assert_parity(got, ref)
    else:
        assert_mixed_parity(got, ref)


TOY_CFG = dict(search_window=(3, 3), estimate_scattering_noise=False)

LAUNCHES = [
    ("fp64", "grid", 256, 1, 64, 1, "lane"),
    ("fp64", "grid", 256, 1, 64, 1, "wave"),
    ("fp64", "persistent", 256, 1, 64, 1, "wave"),
    ("mixed", "grid", 256, 1, 64, 1, "lane"),
    ("mixed", "grid", 64, 1, 64, 1, "wave"),
    ("mixed", "persistent", 256, 1, 64, 1, "wave"),
    ("mixed", "persistent", 64, 2, 128, 16, "wave"),
    ("mixed", "persistent", 128, 1, 256, 64, "wave"),
    ("fp64", "grid", 64, 2, 64, 1, "gather"),
    ("mixed", "grid", 64, 2, 64, 1, "gather"),
    ("mixed", "grid", 256, 1, 64, 1, "gather"),
]


@pytest.mark.parametrize("precision,schedule,block,wps,claim,idle,nav", LAUNCHES)
def test_toy_solenoid_launches(toy7, solenoid, precision, schedule, block, wps, claim, idle, nav):
    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 8192, 44)
    launch = da.LaunchConfig(precision, schedule, block, wps, claim, min_idle=idle, navigation=nav)
    print(f"launch {precision} {schedule} block={block} wps={wps} claim={claim} idle={idle} nav={nav}",
          flush=True)
    got = run_launch(toy7, solenoid, cfg, fp, launch)
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    check(got, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_loopers_persistent(toy7, solenoid, precision):
    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 4096, 45, pt_min=0.1, pt_max=100.0, eta_min=-3.0, eta_max=3.0)
    got = run_launch(toy7, solenoid, cfg, fp, da.LaunchConfig(precision, "persistent"))
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    check(got, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_path_limit_and_noise_persistent(toy7, solenoid, precision):
    # curling tracks reach the path limit; noise estimation widens the mask
    # tolerances of the navigation (ext tolerance)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=True)
    fp = da.generate_tracks("toy", 512, 48, pt_min=0.02, pt_max=0.04, eta_min=-0.3, eta_max=0.3)
    got = run_launch(toy7, solenoid, cfg, fp, da.LaunchConfig(precision, "persistent"))
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    assert np.any(ref["status"] == 2)
    check(got, ref, precision)


def test_telescope_persistent_fp64(telescope):
    cfg = da.PropagationConfig(step_constraint=50.0, estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0)
    fp = da.generate_tracks("beam", 4096, 42)
    got = run_launch(telescope, field, cfg, fp, da.LaunchConfig("fp64", "persistent"))
    ref = oracle_host.propagate(telescope, field, cfg, fp)
    assert np.all(ref["n_surfaces"] == 10)
    assert_parity(got, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_order_permutation_gives_identical_results(toy7, solenoid, precision):
    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 3000, 44)
    base = run_launch(toy7, solenoid, cfg, fp, da.LaunchConfig(precision, "grid", navigation="lane"))
    perm = np.random.default_rng(7).permutation(fp.shape[1]).astype(np.int32)
    order = torch.from_numpy(perm).cuda()
    for schedule, nav in (("grid", "lane"), ("grid", "wave"), ("persistent", "wave"), ("grid", "gather")):
        got = run_launch(toy7, solenoid, cfg, fp, da.LaunchConfig(precision, schedule, order=order,
                                                                   navigation=nav))
        for k in base:
            assert np.array_equal(got[k], base[k]), (schedule, nav, k)


def test_bound_start_persistent(telescope):
    cfg = da.PropagationConfig(step_constraint=50.0, estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 1.0 * da.TESLA, 0.0)
    n = 1024
    rng = np.random.default_rng(3)
    bound = np.zeros((6, n))
    bound[0] = rng.uniform(-50, 50, n)
    bound[1] = rng.uniform(-50, 50, n)
    bound[2] = rng.uniform(-np.pi, np.pi, n)
    bound[3] = rng.uniform(0.02, 0.1, n)
    bound[4] = rng.choice([-1.0, 1.0], n) / rng.uniform(1.0, 10.0, n)
    start = np.full(n, 1, dtype=np.uint32)  # second telescope plane
    for precision in ("fp64", "mixed"):
        got = run_launch(telescope, field, cfg, None, da.LaunchConfig(precision, "persistent"),
                         bound=bound, start=start)
        ref = oracle_host.propagate(telescope, field, cfg, None, bound_params=bound, start_surface=start)
        check(got, ref, precision)


def test_graph_capture_replays(toy7, solenoid):
    """The persistent launch (queue reset kernel + kernel) captured in a HIP
    graph and replayed several times, with eager work between the replays,
    gives the eager result every time."""
    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 4096, 44)
    prop = da.Propagator(cfg, da.LaunchConfig("mixed", "persistent"))
    tracks = torch.from_numpy(fp).cuda()
    eager = prop.propagate(toy7, solenoid, tracks)
    torch.cuda.synchronize()
    ref = {k: getattr(eager, k).cpu().clone() for k in ("bound_params", "bound_cov", "surfaces", "status")}
    out = prop.allocate(fp.shape[1])
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            prop.propagate(toy7, solenoid, tracks, out=out)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        out.status.fill_(255)
        out.bound_params.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, v in ref.items():
            assert torch.equal(getattr(out, k).cpu(), v), (rep, k)
        # eager work between replays
        prop.propagate(toy7, solenoid, tracks[:, :1000].contiguous())
        torch.cuda.synchronize()


def test_native_sort_is_a_permutation_grouped_by_cell():
    from detray_amd.schedule import TrackSorter

    fp = da.generate_tracks("toy", 50000, 44, eta_min=-3.0, eta_max=3.0)
    tracks = torch.from_numpy(fp).cuda()
    sorter = TrackSorter(20, 128, 5.0)
    order = sorter.order(tracks).cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(np.sort(order), np.arange(fp.shape[1]))
    # the cell index is non-decreasing along the order
    dz = np.clip(fp[6], -1 + 1e-12, 1 - 1e-12)
    eta = np.clip(0.5 * np.log((1 + dz) / (1 - dz)), -5.0, 5.0)
    ie = np.minimum(((eta + 5.0) * 20).astype(np.int64), 200)
    phi = np.arctan2(fp[5], fp[4])
    ip = np.minimum(((phi + np.pi) * (128 / (2 * np.pi))).astype(np.int64), 127)
    cell = (ie * 128 + ip)[order]
    # FP32 evaluation on the device may put tracks on a cell boundary into
    # the neighbouring cell
    assert np.mean(np.diff(cell) >= 0) > 0.999


def test_graph_of_sort_and_persistent_propagation_replays(toy7, solenoid):
    """Native sort + persistent propagation captured in one HIP graph and
    replayed several times with eager work in between: every replay gives
    the eager result."""
    from detray_amd.schedule import TrackSorter

    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 4096, 45, pt_min=0.1, pt_max=100.0, eta_min=-3.0, eta_max=3.0)
    tracks = torch.from_numpy(fp).cuda()
    sorter = TrackSorter()
    launch = da.LaunchConfig("mixed", "persistent", min_idle=16)
    prop = da.Propagator(cfg, launch)
    launch.order = sorter.order(tracks)
    eager = prop.propagate(toy7, solenoid, tracks)
    torch.cuda.synchronize()
    ref = {k: getattr(eager, k).cpu().clone() for k in ("bound_params", "bound_cov", "surfaces", "status")}
    out = prop.allocate(fp.shape[1])
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            launch.order = sorter.order(tracks)
            prop.propagate(toy7, solenoid, tracks, out=out)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        out.status.fill_(255)
        out.bound_params.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, v in ref.items():
            assert torch.equal(getattr(out, k).cpu(), v), (rep, k)
        # eager work between the replays (its own propagator: the captured
        # one keeps the order of the full batch)
        da.Propagator(cfg, da.LaunchConfig("mixed", "grid")).propagate(
            toy7, solenoid, tracks[:, :1000].contiguous())
        torch.cuda.synchronize()


@pytest.mark.parametrize("sort_key,nav", [("cells_phi", "gather"), ("cells_phi_fine", "gather"),
                                          ("cells_phi", "lane")])
def test_headline_launch_against_oracle(toy7, solenoid, sort_key, nav):
    """The benchmark's own launch (mixed precision, grid schedule, block 64,
    two waves per SIMD, gathered or per-lane navigation, XCD chunks of 64
    blocks, tracks ordered by the native direction-cell sort) on the first
    131072 tracks of the benchmark batch (toy muons, seed 44, solenoid
    map), against the oracle."""
    from detray_amd.schedule import sort_order

    cfg = da.PropagationConfig(**TOY_CFG)
    fp = da.generate_tracks("toy", 131072, 44)
    tracks = torch.from_numpy(fp).cuda()
    launch = da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, nav, 64)
    launch.order = sort_order(tracks, sort_key)
    print(f"headline launch, sort {sort_key}, navigation {nav}", flush=True)
    res = da.Propagator(cfg, launch).propagate(toy7, solenoid, tracks)
    torch.cuda.synchronize()
    got = {
        "bound_params": res.bound_params.cpu().numpy(),
        "bound_cov": res.bound_cov.cpu().numpy(),
        "path_length": res.path_length.cpu().numpy(),
        "surfaces": res.surfaces.cpu().numpy().view(np.uint32),
        "n_surfaces": res.n_surfaces.cpu().numpy().view(np.uint32),
        "status": res.status.cpu().numpy(),
    }
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    assert_mixed_parity(got, ref)
