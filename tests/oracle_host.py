"""Loader for the oracle (test infrastructure).

The oracle is the FP64 CPU restatement in oracle/oracle.cpp, compiled to
oracle/liboracle.so.  Only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline use it.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import replace as _replace
from pathlib import Path

import numpy as np

from detray_amd import capi

ROOT = Path(__file__).resolve().parent.parent
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

_lib = None


class Trace(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("volume", C.c_void_p), ("surface", C.c_void_p),
                ("n", C.c_void_p)]


def load():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing: run __graft_entry__.build()")
        _lib = C.CDLL(str(ORACLE_LIB))
        _lib.oracle_propagate.restype = C.c_int
        _lib.oracle_field_at.restype = C.c_int
        _lib.oracle_material_probe.restype = C.c_int
    return _lib


def propagate(detector, field, config, free_params, initial_cov=None,
              threads: int = 0, trace_capacity: int = 0, bound_params=None,
              start_surface=None, actors=None) -> dict:
    """Run the oracle on host arrays; returns a dict of numpy arrays.

    Free start: free_params [8, n].  Bound start: bound_params [6, n] and
    start_surface [n] (free_params may be None).  ``actors``: optional
    actors (detray_amd.Actors, target_surface a host array); the step
    tracer's output is out["steps"] (a StepTrace of numpy arrays)."""
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
    cov = None
    if initial_cov is not None:
        cov = np.ascontiguousarray(initial_cov, dtype=np.float64)
    tb = capi.TrackBatch(n, 0, fp.ctypes.data if fp is not None else None,
                         cov.ctypes.data if cov is not None else None,
                         bp.ctypes.data if bp is not None else None,
                         ss.ctypes.data if ss is not None else None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    tr = None
    if trace_capacity:
        out["trace_volume"] = np.zeros((n, trace_capacity), dtype=np.uint32)
        out["trace_surface"] = np.zeros((n, trace_capacity), dtype=np.uint32)
        out["trace_n"] = np.zeros(n, dtype=np.uint32)
        tr = Trace(trace_capacity, out["trace_volume"].ctypes.data,
                   out["trace_surface"].ctypes.data, out["trace_n"].ctypes.data)
    cfg = config.to_c()
    states = None
    if actors is not None and actors.bits():
        if actors.target_surface is not None:
            actors = _replace(actors, target_surface=np.ascontiguousarray(actors.target_surface, dtype=np.uint32))
        out["steps"] = actors.allocate_trace(n, lambda shape, k: np.zeros(shape, dtype={"f8": np.float64,
                                                                                         "i4": np.uint32}[k]))
        states = actors.to_c(out["steps"], lambda a: a.ctypes.data)
        cfg.actors = C.addressof(states)
    rc = lib.oracle_propagate(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg),
                              C.byref(tb), C.byref(rb), C.c_int(threads),
                              C.byref(tr) if tr is not None else None)
    capi.check(rc, "oracle_propagate")
    return out


def propagate_counted(detector, field, config, free_params, threads: int = 0) -> dict:
    """oracle propagate from free starts that also returns n_isect [n], the
    ray-surface intersections each track's navigation computed."""
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
        "n_isect": np.zeros(n, dtype=np.uint32),
    }
    tb = capi.TrackBatch(n, 0, fp.ctypes.data, None, None, None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    lib.oracle_propagate_counted.restype = C.c_int
    cfg = config.to_c()
    capi.check(lib.oracle_propagate_counted(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg), C.byref(tb),
                                            C.byref(rb), C.c_int(threads), C.c_void_p(out["n_isect"].ctypes.data)),
               "oracle_propagate_counted")
    return out


class Hits(C.Structure):
    _fields_ = [("max_hits", C.c_uint32), ("pad", C.c_uint32), ("surface", C.c_void_p), ("loc", C.c_void_p),
                ("n_hits", C.c_void_p)]


def propagate_hits(detector, field, config, bound_params, start_surface, max_hits: int = 32,
                   threads: int = 0) -> dict:
    """Bound start propagation that also returns the updater's loc0/loc1 at
    every sensitive surface crossed: hit_surface [max_hits, n], hit_loc
    [2, max_hits, n], n_hits [n] (the true intersections)."""
    lib = load()
    bp = np.ascontiguousarray(bound_params, dtype=np.float64)
    ss = np.ascontiguousarray(start_surface, dtype=np.uint32)
    n = bp.shape[1]
    out = {
        "bound_params": np.zeros((6, n)),
        "status": np.zeros(n, dtype=np.uint8),
        "n_steps": np.zeros(n, dtype=np.uint32),
        "hit_surface": np.full((max_hits, n), 0xFFFFFFFF, dtype=np.uint32),
        "hit_loc": np.zeros((2, max_hits, n)),
        "n_hits": np.zeros(n, dtype=np.uint32),
    }
    tb = capi.TrackBatch(n, 0, None, None, bp.ctypes.data, ss.ctypes.data)
    rb = capi.ResultBatch(out["bound_params"].ctypes.data, None, None, None, None, None,
                          out["n_steps"].ctypes.data, None, out["status"].ctypes.data)
    h = Hits(max_hits, 0, out["hit_surface"].ctypes.data, out["hit_loc"].ctypes.data, out["n_hits"].ctypes.data)
    cfg = config.to_c()
    cfg.max_surfaces = 0
    lib.oracle_propagate_hits.restype = C.c_int
    capi.check(lib.oracle_propagate_hits(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg), C.byref(tb),
                                         C.byref(rb), C.byref(h), C.c_int(threads), None),
               "oracle_propagate_hits")
    return out


def chi2(rec: dict, meas_surface, meas_loc, meas_n, sigma_loc0: float, sigma_loc1: float) -> np.ndarray:
    """chi2 of the recorded sensitive-surface parameters of each track
    against its measurements: the k-th surface crossed is compared with the
    first not yet used measurement on the same surface (in measurement
    order)."""
    n = rec["n_hits"].shape[0]
    out = np.zeros(n)
    for i in range(n):
        nxt = 0
        nm = int(meas_n[i])
        for k in range(int(rec["n_hits"][i])):
            sf = rec["hit_surface"][k, i]
            for j in range(nxt, nm):
                if meas_surface[j, i] == sf:
                    r0 = rec["hit_loc"][0, k, i] - meas_loc[0, j, i]
                    r1 = rec["hit_loc"][1, k, i] - meas_loc[1, j, i]
                    out[i] += r0 * r0 / sigma_loc0 ** 2 + r1 * r1 / sigma_loc1 ** 2
                    nxt = j + 1
                    break
    return out


def chi2_gradient_fd(detector, field, config, bound_params, start_surface, meas_surface, meas_loc, meas_n,
                     sigma_loc0: float, sigma_loc1: float, rel: float = 1e-6, max_hits: int = 32,
                     params=range(5)) -> np.ndarray:
    """Central finite differences of the oracle's chi2 with respect to the
    bound start parameters, step rel * max(|b_i|, 1) (rel * |q/p| for q/p)."""
    b0 = np.asarray(bound_params, dtype=np.float64)
    g = np.zeros_like(b0)
    for i in params:
        h = rel * (np.abs(b0[i]) if i == 4 else np.maximum(np.abs(b0[i]), 1.0))
        vals = []
        for sgn in (1.0, -1.0):
            b = b0.copy()
            b[i] += sgn * h
            rec = propagate_hits(detector, field, config, b, start_surface, max_hits)
            vals.append(chi2(rec, meas_surface, meas_loc, meas_n, sigma_loc0, sigma_loc1))
        g[i] = (vals[0] - vals[1]) / (2.0 * h)
    return g


def propagate_direct(detector, field, config, free_params, sequence, n_sequence, threads: int = 0,
                     backward: bool = False, bound_params=None, start_surface=None, initial_cov=None) -> dict:
    """Oracle propagation with the direct navigator along the given surface
    sequences [m, n] (n_sequence [n] entries each), forward or backward
    (last entry first).  Free start: free_params [8, n]; bound start:
    bound_params [6, n], start_surface [n] and optionally initial_cov
    [21, n] (free_params may be None)."""
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
                         bp.ctypes.data if bp is not None else None,
                         ss.ctypes.data if ss is not None else None)
    rb = capi.ResultBatch(*[out[k].ctypes.data for k in (
        "bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces",
        "n_steps", "n_trials", "status")])
    lib.oracle_propagate_direct.restype = C.c_int
    cfg = config.to_c()
    direction = capi.DTR_DIRECTION_BACKWARD if backward else capi.DTR_DIRECTION_FORWARD
    capi.check(lib.oracle_propagate_direct(C.byref(detector.desc), C.byref(field.desc), C.byref(cfg), C.byref(tb),
                                           C.c_void_p(seq.ctypes.data), C.c_void_p(nseq.ctypes.data),
                                           C.c_uint32(seq.shape[0]), C.c_uint32(direction), C.byref(rb),
                                           C.c_int(threads)),
               "oracle_propagate_direct")
    return out


def field_at(field, xyz: np.ndarray) -> np.ndarray:
    lib = load()
    pts = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pts)
    capi.check(lib.oracle_field_at(C.byref(field.desc), C.c_uint32(len(pts)),
                                   pts.ctypes.data_as(C.POINTER(C.c_double)),
                                   out.ctypes.data_as(C.POINTER(C.c_double))), "oracle_field_at")
    return out


def surface_candidates(detector, sf: int, ro, rd, tol: float, ext_tol: float = 0.0, overstep: float = 0.0):
    """The navigator's candidate intersections with one surface:
    list of (path, global point, status 0 inside / 1 edge)."""
    lib = load()
    out = (C.c_double * 10)()
    n = lib.oracle_surface_candidates(C.byref(detector.desc), C.c_uint32(sf), (C.c_double * 3)(*ro),
                                      (C.c_double * 3)(*rd), C.c_double(tol), C.c_double(ext_tol),
                                      C.c_double(overstep), out)
    return [(out[5 * k], np.array(out[5 * k + 1:5 * k + 4]), int(out[5 * k + 4])) for k in range(n)]


def line_intersect(ro, rd, transform, mask, tol: float):
    """Ray-wire intersection of the oracle: (valid, path, loc0, loc1, inside)."""
    lib = load()
    out = (C.c_double * 4)()
    ok = lib.oracle_line_intersect((C.c_double * 3)(*ro), (C.c_double * 3)(*rd), C.byref(transform),
                                   C.byref(mask), C.c_double(tol), out)
    return bool(ok), out[0], out[1], out[2], bool(out[3])


def material_probe(slab, qop: float, path: float, x_over_x0: float) -> np.ndarray:
    lib = load()
    out = (C.c_double * 3)()
    capi.check(lib.oracle_material_probe(C.byref(slab), C.c_double(qop), C.c_double(path),
                                         C.c_double(x_over_x0), out), "oracle_material_probe")
    return np.array(out[:])


def eloss_probe(mat, qop: float) -> np.ndarray:
    """{-dE/dx, d(qop)/ds, d(d(qop)/ds)/dqop} of the muon hypothesis."""
    lib = load()
    out = (C.c_double * 3)()
    capi.check(lib.oracle_eloss_probe(C.byref(mat), C.c_double(qop), out), "oracle_eloss_probe")
    return np.array(out[:])


def rk_steps(field, vol_mat, config, free_params, ds: float, n_steps: int):
    """The oracle's stepper alone: n_steps steps of ds (reset each time)
    in the volume material vol_mat (None: vacuum).  Returns (rows
    [n_steps + 1, 5] of qop, dqopds, d2qopdsdqop, path length, J(qop, qop);
    the final free Jacobian [8, 8]; the final free parameters [8])."""
    lib = load()
    fp = np.ascontiguousarray(free_params, dtype=np.float64)
    out = np.zeros((n_steps + 1, 5))
    jac = np.zeros(72)
    cfg = config.to_c()
    capi.check(lib.oracle_rk_steps(C.byref(field.desc), C.byref(vol_mat) if vol_mat is not None else None,
                                   C.byref(cfg), C.c_void_p(fp.ctypes.data), C.c_double(ds),
                                   C.c_uint32(n_steps), C.c_void_p(out.ctypes.data),
                                   C.c_void_p(jac.ctypes.data)), "oracle_rk_steps")
    return out, jac[:64].reshape(8, 8), jac[64:]


def default_threads() -> int:
    return int(os.environ.get("ORACLE_THREADS", "0"))
