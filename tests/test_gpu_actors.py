"""Optional actors on the device against the oracle: momentum and target
aborters and the step tracer, on the plain FP64 kernel, the batched
kernels (FP64 and mixed, per-lane, gathered, wave-pooled and persistent)
and the direct navigator.

The step tracer carries the reference's own host/device contract
(tests/include/detray/test/device/propagator_test.hpp:171-232: per recorded
step, path length and position within 1e-4 of the path length, transport
Jacobian elements within 1e-4 times the path length); the FP64 kernels are
held to 1e-9 instead.  Statuses and surface sequences are identical."""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

TOY = dict(search_window=(3, 3), estimate_scattering_noise=False)


def _trace_host(tr):
    return {k: (getattr(tr, k).cpu().numpy() if getattr(tr, k) is not None else None)
            for k in ("n", "surface", "step_size", "path_length", "n_trials", "free_params", "bound_params",
                      "jacobian")}


def _check(got, ref, tol):
    assert np.array_equal(got["status"], ref["status"])
    assert np.array_equal(got["n_surfaces"], ref["n_surfaces"].astype(np.int32))
    gt, rt = _trace_host(got["trace"]), ref["steps"]
    assert np.array_equal(gt["n"].view(np.uint32), rt.n)
    cap = rt.surface.shape[0]
    k = np.arange(cap)[:, None]
    valid = k < np.minimum(rt.n, cap)[None, :]
    assert np.array_equal(np.where(valid, gt["surface"].view(np.uint32), 0), np.where(valid, rt.surface, 0))
    path = np.maximum(np.abs(rt.path_length), 1.0)
    assert np.max(np.abs(gt["path_length"] - rt.path_length)[valid] / path[valid], initial=0.0) <= tol
    dpos = np.linalg.norm(gt["free_params"][0:3] - rt.free_params[0:3], axis=0)
    assert np.max((dpos / path)[valid], initial=0.0) <= tol
    dj = np.abs(gt["jacobian"] - rt.jacobian).max(axis=0)
    assert np.max((dj / path)[valid], initial=0.0) <= tol


def _run(prop, det, field, fp, actors):
    res = prop.propagate(det, field, torch.from_numpy(fp).cuda(), actors=actors)
    torch.cuda.synchronize()
    return {"status": res.status.cpu().numpy(), "n_surfaces": res.n_surfaces.cpu().numpy(), "trace": res.trace,
            "bound_params": res.bound_params.cpu().numpy()}


LAUNCHES = [None,
            da.LaunchConfig("fp64", "grid", 64, 2, 64, None, 1, "gather"),
            da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "gather"),
            da.LaunchConfig("mixed", "grid", 256, 1, 64, None, 1, "lane"),
            da.LaunchConfig("mixed", "grid", 64, 2, 64, None, 1, "wave"),
            da.LaunchConfig("fp64", "persistent", 256, 2, 64, None, 16, "wave")]


@pytest.mark.parametrize("launch", LAUNCHES, ids=lambda l: "plain" if l is None else
                         f"{l.precision}-{l.schedule}-{l.navigation}-{l.block_size}")
def test_actors_against_oracle(toy7, solenoid, launch):
    cfg = da.PropagationConfig(**TOY)
    fp = da.generate_tracks("toy", 4096, 44, pt_min=0.3, pt_max=5.0)
    n = fp.shape[1]
    plain_ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    j = np.minimum(3, np.maximum(plain_ref["n_surfaces"].astype(np.int64) - 1, 0))
    target = np.where(plain_ref["n_surfaces"] > 0, plain_ref["surfaces"][j, np.arange(n)], 0).astype(np.uint32)
    target[np.arange(n) % 7 == 0] = toy7.n_surfaces - 1  # never crossed
    host = da.Actors(momentum_aborter=True, min_pT=0.8, min_p=1.2, target_surface=target, step_tracer=16,
                     trace_jacobian=True)
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp, actors=host)
    assert (ref["status"] == 4).sum() > 100 and (ref["status"] == 5).sum() > 100
    dev = da.Actors(momentum_aborter=True, min_pT=0.8, min_p=1.2,
                    target_surface=torch.from_numpy(target.view(np.int32)).cuda(), step_tracer=16,
                    trace_jacobian=True)
    print(f"actors, launch {launch}", flush=True)
    got = _run(da.Propagator(cfg, launch), toy7, solenoid, fp, dev)
    mixed = launch is not None and launch.precision == "mixed"
    _check(got, ref, 1e-4 if mixed else 1e-9)


def test_direct_navigator_with_actors(toy7, solenoid):
    cfg = da.PropagationConfig(**TOY)
    fp = da.generate_tracks("toy", 2048, 46)
    base = da.Propagator(cfg).propagate(toy7, solenoid, torch.from_numpy(fp).cuda())
    torch.cuda.synchronize()
    host = da.Actors(step_tracer=16, trace_jacobian=True)
    res = da.Propagator(cfg, da.LaunchConfig("fp64", "grid", 64, 2, 64, None, 1, "lane")).propagate_direct(
        toy7, solenoid, base.surfaces, base.n_surfaces, torch.from_numpy(fp).cuda(), actors=host)
    torch.cuda.synchronize()
    seq = base.surfaces.cpu().numpy().view(np.uint32)
    ref = oracle_host.propagate_direct(toy7, solenoid, cfg, fp, seq, base.n_surfaces.cpu().numpy().view(np.uint32))
    tr = _trace_host(res.trace)
    # the tracer records the replayed sequence, and the direct navigator's
    # final parameters are the oracle's
    n = fp.shape[1]
    assert np.array_equal(tr["n"], res.n_surfaces.cpu().numpy())
    for i in range(0, n, 97):
        m = int(tr["n"][i])
        assert np.array_equal(tr["surface"][:m, i].view(np.uint32), seq[:m, i])
    scale = np.maximum(np.abs(ref["bound_params"]), 1.0)
    assert np.max(np.abs(res.bound_params.cpu().numpy() - ref["bound_params"]) / scale) <= 1e-9
