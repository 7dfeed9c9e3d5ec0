"""Loader for the host build of the device engine (test infrastructure).

tests/native/engine_host.cpp compiles detray_amd/csrc/device/engine.hpp,
the code the batched kernels run, for the CPU (hipcc --offload-host-only),
so that CPU tests can check it: the FP32 grid prefilter against the exact
mask check, and the per-lane propagation loop against the oracle."""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from detray_amd import capi

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "tests" / "native" / "libengine_host.so"

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run __graft_entry__.build()")
        _lib = C.CDLL(str(LIB))
        _lib.dtr_test_prefilter.restype = C.c_int
        _lib.dtr_test_engine_propagate.restype = C.c_int
    return _lib


def prefilter(detector, config, rays, ext_tol: float = 0.0):
    """rays [n, 6] (origin, unit direction).  Returns (pairs evaluated,
    pairs the exact check accepts, accepted pairs the prefilter rejects,
    (ray, surface) of the first such miss)."""
    lib = load()
    r = np.ascontiguousarray(rays, dtype=np.float64)
    out = np.zeros(3, dtype=np.uint64)
    miss = np.zeros(2, dtype=np.uint32)
    cfg = config.to_c()
    capi.check(lib.dtr_test_prefilter(C.byref(detector.desc), C.byref(cfg), C.c_void_p(r.ctypes.data),
                                      C.c_uint32(r.shape[0]), C.c_double(ext_tol),
                                      C.c_void_p(out.ctypes.data), C.c_void_p(miss.ctypes.data)),
               "dtr_test_prefilter")
    return int(out[0]), int(out[1]), int(out[2]), (int(miss[0]), int(miss[1]))


def propagate(detector, field, config, free_params, mixed: bool = False, initial_cov=None, bound_params=None,
              start_surface=None, plain: bool = False) -> dict:
    """The engine's lane loop (propagate_fast) over host arrays, FP64 or
    mixed policy; with plain=True the plain FP64 kernel's per-track loop
    (propagate_track); the result arrays of dtr_propagate_ex."""
    lib = load()
    fp = None if free_params is None else np.ascontiguousarray(free_params, dtype=np.float64)
    bp = None if bound_params is None else np.ascontiguousarray(bound_params, dtype=np.float64)
    ss = None if start_surface is None else np.ascontiguousarray(start_surface, dtype=np.uint32)
    n = fp.shape[1] if fp is not None else bp.shape[1]
    ms = max(int(config.max_surfaces), 1)
    out = {
        "bound_params": np.zeros((6, n)),
        "bound_cov": np.zeros((21, n)),
        "free_params": np.zeros((8, n)),
        "path_length": np.zeros(n),
        "surfaces": np.full((ms, n), 0xFFFFFFFF, dtype=np.uint32),
        "n_surfaces": np.zeros(n, dtype=np.uint32),
        "n_steps": np.zeros(n, dtype=np.uint32),
        "n_trials": np.zeros(n, dtype=np.uint32),
        "status": np.zeros(n, dtype=np.uint8),
    }
    cov = None if initial_cov is None else np.ascontiguousarray(initial_cov, dtype=np.float64)
    tb = capi.TrackBatch(n, 0, fp.ctypes.data if fp is not None else None,
                         cov.ctypes.data if cov is not None else None,
                         bp.ctypes.data if bp is not None else None, ss.ctypes.data if ss is not None else None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    cfg = config.to_c()
    if plain:
        capi.check(lib.dtr_test_plain_propagate(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg),
                                                C.byref(tb), C.byref(rb)), "dtr_test_plain_propagate")
        return out
    capi.check(lib.dtr_test_engine_propagate(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg),
                                             C.byref(tb), C.byref(rb), C.c_int(1 if mixed else 0)),
               "dtr_test_engine_propagate")
    return out


def propagate_direct(detector, field, config, free_params, sequence, n_sequence, mixed: bool = False,
                     backward: bool = False, bound_params=None, start_surface=None, initial_cov=None) -> dict:
    """The engine's direct-navigator loop (propagate_direct) over host
    arrays, forward or backward, FP64 or mixed policy."""
    lib = load()
    fp = None if free_params is None else np.ascontiguousarray(free_params, dtype=np.float64)
    bp = None if bound_params is None else np.ascontiguousarray(bound_params, dtype=np.float64)
    ss = None if start_surface is None else np.ascontiguousarray(start_surface, dtype=np.uint32)
    cov = None if initial_cov is None else np.ascontiguousarray(initial_cov, dtype=np.float64)
    seq = np.ascontiguousarray(sequence, dtype=np.uint32)
    nseq = np.ascontiguousarray(n_sequence, dtype=np.uint32)
    n = fp.shape[1] if fp is not None else bp.shape[1]
    ms = max(int(config.max_surfaces), 1)
    out = {
        "bound_params": np.zeros((6, n)),
        "bound_cov": np.zeros((21, n)),
        "free_params": np.zeros((8, n)),
        "path_length": np.zeros(n),
        "surfaces": np.full((ms, n), 0xFFFFFFFF, dtype=np.uint32),
        "n_surfaces": np.zeros(n, dtype=np.uint32),
        "n_steps": np.zeros(n, dtype=np.uint32),
        "n_trials": np.zeros(n, dtype=np.uint32),
        "status": np.zeros(n, dtype=np.uint8),
    }
    tb = capi.TrackBatch(n, 0, fp.ctypes.data if fp is not None else None,
                         cov.ctypes.data if cov is not None else None,
                         bp.ctypes.data if bp is not None else None, ss.ctypes.data if ss is not None else None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    direction = capi.DTR_DIRECTION_BACKWARD if backward else capi.DTR_DIRECTION_FORWARD
    sq = capi.SurfaceSequence(int(seq.shape[0]), direction, seq.ctypes.data, nseq.ctypes.data)
    cfg = config.to_c()
    capi.check(lib.dtr_test_engine_propagate_direct(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg),
                                                    C.byref(tb), C.byref(sq), C.byref(rb),
                                                    C.c_int(1 if mixed else 0)),
               "dtr_test_engine_propagate_direct")
    return out


def adjoint(detector, field, config, adjoint_config, bound_params, start_surface, hit_surface=None, hit_loc=None,
            n_hits=None, transport_gradient: bool = False, backward: bool = True) -> dict:
    """Both passes of the adjoint (adjoint_record, adjoint_backward) over
    host arrays: chi2 [n], status [n], grad [6, n] (backward) and the
    Jacobian-chain gradient grad_transport [6, n] (transport_gradient)."""
    lib = load()
    bp = np.ascontiguousarray(bound_params, dtype=np.float64)
    ss = np.ascontiguousarray(start_surface, dtype=np.uint32)
    n = bp.shape[1]
    out = {"chi2": np.zeros(n), "status": np.zeros(n, dtype=np.uint32),
           "grad": np.zeros((6, n)) if backward else None,
           "grad_transport": np.zeros((6, n)) if transport_gradient else None}
    hits = None
    if hit_surface is not None:
        hs = np.ascontiguousarray(hit_surface, dtype=np.uint32)
        hl = np.ascontiguousarray(hit_loc, dtype=np.float64)
        hn = np.ascontiguousarray(n_hits, dtype=np.uint32)
        out["_keep"] = (hs, hl, hn)
        hits = capi.Hits(int(hs.shape[0]), 0, hs.ctypes.data, hl.ctypes.data, hn.ctypes.data)
    cfg = config.to_c()
    acfg = adjoint_config.to_c()
    ptr = lambda a: None if a is None else C.c_void_p(a.ctypes.data)  # noqa: E731
    lib.dtr_test_adjoint.restype = C.c_int
    capi.check(lib.dtr_test_adjoint(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg), C.byref(acfg),
                                    C.c_uint32(n), ptr(bp), ptr(ss), C.byref(hits) if hits is not None else None,
                                    ptr(out["chi2"]), ptr(out["grad_transport"]), ptr(out["grad"]),
                                    ptr(out["status"])),
               "dtr_test_adjoint")
    out.pop("_keep", None)
    return out


def propagate_resumable(detector, field, config, free_params, max_steps: int, mixed: bool = False,
                        max_calls: int = 100000) -> tuple:
    """The engine's resumable lane loop (propagate_resumable) over host
    arrays, called with a budget of ``max_steps`` steps until no track is
    paused; returns (results as propagate(), number of calls, number of
    tracks paused by the first call)."""
    lib = load()
    fp = np.ascontiguousarray(free_params, dtype=np.float64)
    n = fp.shape[1]
    ms = max(int(config.max_surfaces), 1)
    out = {
        "bound_params": np.zeros((6, n)),
        "bound_cov": np.zeros((21, n)),
        "free_params": np.zeros((8, n)),
        "path_length": np.zeros(n),
        "surfaces": np.full((ms, n), 0xFFFFFFFF, dtype=np.uint32),
        "n_surfaces": np.zeros(n, dtype=np.uint32),
        "n_steps": np.zeros(n, dtype=np.uint32),
        "n_trials": np.zeros(n, dtype=np.uint32),
        "status": np.zeros(n, dtype=np.uint8),
    }
    state = np.zeros(int(capi.load().dtr_propagation_state_size(n)) // 8 + 1, dtype=np.float64)
    tb = capi.TrackBatch(n, 0, fp.ctypes.data, None, None, None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    cfg = config.to_c()
    lib.dtr_test_engine_propagate_resumable.restype = C.c_int
    calls, first_paused = 0, 0
    while calls < max_calls:
        capi.check(lib.dtr_test_engine_propagate_resumable(
            C.byref(detector.desc), C.byref(field.desc), C.byref(cfg), C.byref(tb), C.byref(rb),
            C.c_void_p(state.ctypes.data), C.c_uint32(max_steps), C.c_uint32(1 if calls else 0),
            C.c_int(1 if mixed else 0)), "dtr_test_engine_propagate_resumable")
        paused = int((out["status"] == capi.TRACK_PAUSED).sum())
        if calls == 0:
            first_paused = paused
        calls += 1
        if paused == 0:
            break
    return out, calls, first_paused
