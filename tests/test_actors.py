"""The optional actors of the chain on the oracle (CPU): the step tracer
(step_tracer.hpp, observer of the parameter transporter), the momentum
aborter (aborters.hpp:80-143) and the target aborter (aborters.hpp:146-175).

The aborters are pinned by their definition against runs without them: a
run with the step tracer gives each track's momentum after every surface's
material interaction and its surface sequence; the momentum aborter must
stop a track at the first surface where pT or p has fallen to its limit
(or at the start if it starts there), the target aborter at its target
surface, and both must leave every other track untouched."""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host

TRACK_EXIT, TRACK_MOMENTUM, TRACK_TARGET = 0, 4, 5


@pytest.fixture(scope="module")
def toy_const(toy7):
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    return toy7, field, cfg


def _traced(det, field, cfg, fp, cap=32, **kw):
    return oracle_host.propagate(det, field, cfg, fp, actors=da.Actors(step_tracer=cap, trace_jacobian=True, **kw))


def test_step_tracer_records_every_parameter_update(toy_const):
    det, field, cfg = toy_const
    fp = da.generate_tracks("toy", 2000, 43)
    plain = oracle_host.propagate(det, field, cfg, fp)
    r = _traced(det, field, cfg, fp)
    st = r["steps"]
    # the tracer observes the updater and changes nothing
    for k in ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "status"):
        assert np.array_equal(r[k], plain[k]), k
    # one entry per sensitive or material surface: the sequencer's surfaces
    assert np.array_equal(st.n, r["n_surfaces"])
    n = fp.shape[1]
    for i in range(n):
        m = int(st.n[i])
        assert np.array_equal(st.surface[:m, i], r["surfaces"][:m, i])
        assert np.all(np.diff(st.path_length[:m, i]) > 0.0)
    # the last entry holds the final bound parameters (after its material)
    has = st.n > 0
    last = (st.n[has] - 1).astype(np.int64)
    cols = np.flatnonzero(has)
    assert np.array_equal(st.bound_params[:, last, cols], r["bound_params"][:, has])
    # recorded positions lie on their surfaces: planar modules at local z = 0
    tr = det.table("transforms")
    sfs = det.table("surfaces")
    worst = 0.0
    for i in cols[:500]:
        for k in range(int(st.n[i])):
            s = sfs[int(st.surface[k, i])]
            if s.shape in (0, 1):  # rectangle, trapezoid
                t = tr[s.transform]
                p = st.free_params[0:3, k, i] - np.array(t.t)
                normal = np.array([t.r[2], t.r[5], t.r[8]])
                worst = max(worst, abs(float(p @ normal)))
    assert worst < 1.5e-3  # within the navigator's on-surface tolerance (1 um) and the step's landing
    # the transport Jacobian: reset to the identity by every transport after
    # the first surface; at the first surface the accumulated one (not reset:
    # the departure is the free start)
    eye = np.eye(8).reshape(64)
    for i in cols[:500]:
        m = int(st.n[i])
        assert not np.allclose(st.jacobian[:, 0, i], eye)
        for k in range(1, m):
            assert np.array_equal(st.jacobian[:, k, i], eye)


def test_step_tracer_capacity_counts_beyond(toy_const):
    det, field, cfg = toy_const
    fp = da.generate_tracks("toy", 500, 43)
    full = _traced(det, field, cfg, fp, cap=32)["steps"]
    short = _traced(det, field, cfg, fp, cap=2)["steps"]
    assert np.array_equal(short.n, full.n)
    assert np.array_equal(short.surface, full.surface[:2])
    assert np.array_equal(short.free_params, full.free_params[:, :2])


@pytest.mark.parametrize("which", ["pT", "p"])
def test_momentum_aborter_stops_at_the_first_surface_below_the_limit(toy_const, which):
    det, field, cfg = toy_const
    fp = da.generate_tracks("toy", 3000, 44, pt_min=0.3, pt_max=3.0)
    ref = _traced(det, field, cfg, fp)
    st = ref["steps"]
    qop0 = fp[7]
    pt0 = np.abs(np.hypot(fp[4], fp[5]) / qop0)
    p0 = np.abs(1.0 / qop0)
    limit = 1.0 if which == "pT" else 1.5
    actors = da.Actors(momentum_aborter=True, min_pT=limit if which == "pT" else 0.0,
                       min_p=limit if which == "p" else 0.0, step_tracer=32)
    got = oracle_host.propagate(det, field, cfg, fp, actors=actors)
    n = fp.shape[1]
    n_stop = n_keep = 0
    for i in range(n):
        m = int(st.n[i])
        qop = st.bound_params[4, :m, i]
        th = st.bound_params[3, :m, i]
        after = np.abs(np.sin(th) / qop) if which == "pT" else np.abs(1.0 / qop)
        start = pt0[i] if which == "pT" else p0[i]
        if start <= limit:
            # at the first actor call, before any step
            assert got["status"][i] == TRACK_MOMENTUM and got["path_length"][i] == 0.0
            assert got["n_surfaces"][i] == 0
            n_stop += 1
            continue
        below = np.flatnonzero(after <= limit)
        if below.size:
            k = int(below[0])
            assert got["status"][i] == TRACK_MOMENTUM, i
            assert got["n_surfaces"][i] == k + 1
            assert np.array_equal(got["surfaces"][:k + 1, i], ref["surfaces"][:k + 1, i])
            assert np.array_equal(got["bound_params"][:, i], st.bound_params[:, k, i])
            assert got["path_length"][i] == st.path_length[k, i]
            n_stop += 1
        else:
            for key in ("bound_params", "path_length", "n_surfaces", "status"):
                assert np.array_equal(got[key][..., i], ref[key][..., i]), (i, key)
            n_keep += 1
    assert n_stop > 100 and n_keep > 100


def test_target_aborter_stops_on_the_target(toy_const):
    det, field, cfg = toy_const
    fp = da.generate_tracks("toy", 3000, 45)
    ref = _traced(det, field, cfg, fp)
    st = ref["steps"]
    n = fp.shape[1]
    j = np.minimum(2, np.maximum(ref["n_surfaces"].astype(np.int64) - 1, 0))
    target = np.where(ref["n_surfaces"] > 0, ref["surfaces"][j, np.arange(n)], 0).astype(np.uint32)
    # a few tracks get a target they never cross: they must run to the end
    far = np.arange(n) % 10 == 0
    target[far] = det.n_surfaces - 1
    got = oracle_host.propagate(det, field, cfg, fp, actors=da.Actors(target_surface=target))
    for i in range(n):
        crosses = target[i] in ref["surfaces"][: int(ref["n_surfaces"][i]), i]
        if not crosses:
            assert got["status"][i] == ref["status"][i] and got["n_surfaces"][i] == ref["n_surfaces"][i]
            continue
        k = int(np.flatnonzero(ref["surfaces"][: int(ref["n_surfaces"][i]), i] == target[i])[0])
        assert got["status"][i] == TRACK_TARGET
        assert got["n_surfaces"][i] == k + 1
        assert np.array_equal(got["bound_params"][:, i], st.bound_params[:, k, i])
        assert got["path_length"][i] == st.path_length[k, i]


def test_actor_states_are_checked():
    det = da.Detector.telescope([100.0, 200.0], 50.0, 50.0)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    fp = da.generate_tracks("beam", 8, 42)
    from detray_amd import capi

    a = da.Actors(step_tracer=4)
    # a step tracer without its count array is an error
    states = a.to_c(None, lambda x: 0)
    states.trace.n = None
    cfg = da.PropagationConfig().to_c()
    import ctypes as C

    cfg.actors = C.addressof(states)
    out = {k: np.zeros(s) for k, s in (("b", (6, 8)),)}
    tb = capi.TrackBatch(8, 0, np.ascontiguousarray(fp).ctypes.data, None, None, None)
    rb = capi.ResultBatch(out["b"].ctypes.data, None, None, None, None, None, None, None, None)
    rc = oracle_host.load().oracle_propagate(C.byref(det.desc), C.byref(field.desc), C.byref(cfg), C.byref(tb),
                                             C.byref(rb), C.c_int(1), None)
    assert rc == 1
