"""Track schedules and the launch autotuner for the batched kernels.

A wave of 64 lanes runs 64 tracks in lock step; tracks that cross the same
volumes and grid bins at the same time keep the lanes on the same code path
and the same detector records.  :func:`sort_order` builds that schedule on
the device (a permutation of the batch, passed as ``LaunchConfig.order``),
and :func:`autotune` picks the block size, register budget (waves per SIMD),
schedule, navigation mode, persistent claim size, XCD chunk and sort key by
timing candidates on a sample of the batch.
"""

from __future__ import annotations

import json
import math
import os
import time
from dataclasses import asdict
from pathlib import Path
from typing import Iterable, Optional

from .propagator import LaunchConfig, Propagator

SORT_KEYS = ("none", "cells", "cells_coarse", "cells_fine", "cells_phi", "cells_phi_fine", "eta_phi",
             "eta_phi_coarse", "phi_eta", "eta_phi_pt")

# direction-cell granularities of the native sort: (eta bins per unit, phi bins)
CELL_KEYS = {"cells": (20, 128), "cells_coarse": (10, 64), "cells_fine": (40, 256), "cells_phi": (10, 8192),
             "cells_phi_fine": (10, 32768)}


class TrackSorter:
    """Native counting sort of a batch by direction cell
    (``dtr_sort_tracks``): the order and workspace buffers are allocated once
    per batch size, and a call only launches kernels, so it can be captured
    in a HIP graph together with the propagation."""

    def __init__(self, eta_bins_per_unit: int = 20, phi_bins: int = 128, eta_max: float = 5.0):
        from . import capi

        self._capi = capi
        self.cfg = capi.SortConfig()
        capi.load().dtr_default_sort(self.cfg)
        self.cfg.eta_bins_per_unit = int(eta_bins_per_unit)
        self.cfg.phi_bins = int(phi_bins)
        self.cfg.eta_max = float(eta_max)
        self._buf = {}

    def order(self, tracks):
        import ctypes as C

        import torch

        lib = self._capi.load()
        n = int(tracks.shape[1])
        key = (str(tracks.device), n)
        if key not in self._buf:
            ws = int(lib.dtr_sort_workspace_size(C.byref(self.cfg)))
            self._buf[key] = (torch.empty(n, dtype=torch.int32, device=tracks.device),
                              torch.empty((ws + 3) // 4, dtype=torch.int32, device=tracks.device), ws)
        order, work, ws = self._buf[key]
        stream = torch.cuda.current_stream(tracks.device).cuda_stream
        self._capi.check(lib.dtr_sort_tracks(C.byref(self.cfg), C.c_void_p(tracks.data_ptr()), n,
                                             C.c_void_p(order.data_ptr()), C.c_void_p(work.data_ptr()),
                                             ws, C.c_void_p(stream)), "dtr_sort_tracks")
        return order


_SORTERS = {}


def sort_order(tracks, key: str = "cells", eta_bins_per_unit: float = 20.0):
    """Permutation (int32, on the tracks' device) that groups tracks by
    direction.  ``cells*``: the native counting sort by (eta, phi) cell
    (graph-capturable).  ``eta_phi``: torch argsort by pseudorapidity bins of
    width 1/eta_bins_per_unit, then phi; ``phi_eta`` swaps the two;
    ``eta_phi_pt`` adds the transverse momentum inside phi slices.
    ``tracks`` is the free-parameter tensor [8, n].  Returns None for
    ``key == "none"``."""
    import torch

    if key == "none":
        return None
    if key in CELL_KEYS:
        if key not in _SORTERS:
            _SORTERS[key] = TrackSorter(*CELL_KEYS[key])
        return _SORTERS[key].order(tracks)
    if key == "eta_phi_coarse":
        key, eta_bins_per_unit = "eta_phi", eta_bins_per_unit / 2.0
    dx, dy, dz = tracks[4], tracks[5], tracks[6]
    dz = dz.clamp(-1.0 + 1e-12, 1.0 - 1e-12)
    eta = 0.5 * torch.log((1.0 + dz) / (1.0 - dz))
    phi = torch.atan2(dy, dx)
    eta_bin = torch.round(eta * eta_bins_per_unit).to(torch.int64) + 4096
    phi_q = ((phi + math.pi) * (2 ** 20 / (2 * math.pi))).to(torch.int64).clamp(0, 2 ** 20 - 1)
    if key == "eta_phi":
        k = (eta_bin << 21) | phi_q
    elif key == "phi_eta":
        phi_bin = (phi_q >> 14)  # 64 phi slices
        eta_q = ((eta + 8.0) * (2 ** 16)).to(torch.int64).clamp(0, 2 ** 20 - 1)
        k = (phi_bin << 21) | eta_q
    elif key == "eta_phi_pt":
        pt = torch.hypot(dx, dy) / tracks[7].abs()
        pt_q = (torch.log(pt.clamp(min=1e-3)) * 8.0 + 64.0).to(torch.int64).clamp(0, 127)
        k = (eta_bin << 28) | ((phi_q >> 14) << 7) | pt_q
    else:
        raise ValueError(f"unknown sort key {key!r}")
    return torch.argsort(k).to(torch.int32)


def candidate_launches(precision: str = "mixed") -> Iterable[tuple]:
    """The autotuner's search space: (schedule, navigation, block, waves,
    claim, min_idle, sort key, xcd chunk)."""
    for key in ("cells", "cells_phi", "cells_phi_fine", "eta_phi", "none"):
        for block in (64, 128, 256):
            for waves in (1, 2):
                yield ("grid", "gather", block, waves, 64, 1, key, 64)
        yield ("grid", "gather", 64, 2, 64, 1, key, 0)
        yield ("grid", "gather", 64, 2, 64, 1, key, 256)
        yield ("grid", "lane", 64, 2, 64, 1, key, 64)
        yield ("grid", "wave", 64, 2, 64, 1, key, 0)
        for claim in (64, 256):
            for idle in (16, 48):
                yield ("persistent", "wave", 256, 2, claim, idle, key, 0)


def _cache_path() -> Path:
    base = os.environ.get("DETRAY_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "detray_amd"))
    return Path(base) / "autotune.json"


def autotune(detector, field, config, tracks, precision: str = "mixed", sample: int = 262144,
             reps: int = 2, candidates: Optional[Iterable[tuple]] = None, use_cache: bool = True,
             verbose: bool = False, tag: str = ""):
    """Time every candidate launch on the first ``sample`` tracks (sort
    included) and return ``(LaunchConfig without order, sort key, table)``.
    Results are cached per (detector, field, precision, sample size, device,
    ``tag`` naming the track population) in
    ``$DETRAY_AMD_CACHE/autotune.json``."""
    import torch

    n = int(tracks.shape[1])
    m = min(sample, n)
    dev_name = torch.cuda.get_device_name(tracks.device)
    cache_key = f"{detector.name}|{field.name}|{precision}|{m}|{dev_name}|{tag}"
    cache = {}
    path = _cache_path()
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if cache_key in cache:
            e = cache[cache_key]
            return LaunchConfig(**e["launch"]), e["sort"], e["table"]
    sub = tracks[:, :m].contiguous()
    table = []
    best = None
    for cand in (candidates or candidate_launches(precision)):
        schedule, nav, block, waves, claim, idle, key = cand[:7]
        xcd = cand[7] if len(cand) > 7 else 64
        launch = LaunchConfig(precision, schedule, block, waves, claim, None, idle, nav, xcd)
        prop = Propagator(config, launch)
        out = prop.allocate(m, sub.device)
        launch.order = sort_order(sub, key)
        prop.propagate(detector, field, sub, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            launch.order = sort_order(sub, key)
            prop.propagate(detector, field, sub, out=out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        row = {"schedule": schedule, "navigation": nav, "block": block, "waves_per_simd": waves,
               "tracks_per_claim": claim, "min_idle": idle, "sort": key, "xcd_chunk": xcd, "ms": dt * 1e3}
        table.append(row)
        if verbose:
            print(f"[autotune] {row}", flush=True)
        if best is None or dt < best[0]:
            best = (dt, launch, key)
    _, launch, key = best
    launch.order = None
    if use_cache:
        cache[cache_key] = {"launch": {k: v for k, v in asdict(launch).items() if k != "order"},
                            "sort": key, "table": table}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return launch, key, table
