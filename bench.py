"""Benchmark: tracks/s of detray's propagation hot path on one MI355X.

Workload (one "step" = one propagate() pass over the whole batch, the track
sort of the schedule included):
  toy detector (4 barrel layers, 7 endcap discs per side, beampipe, grids,
  homogeneous Si material), analytic finite-solenoid field (2 T centre,
  R = 1.2 m, half length 3 m) sampled on a 5 cm fp32 xyz grid with
  trilinear interpolation; 1M muons, eta ~ U[-2.5, 2.5], phi ~ U,
  pT log-uniform [1, 100] GeV, vertex sigma_xy = 10 um, sigma_z = 50 mm;
  search window 3x3, covariance transport, parameter_updater with
  pointwise material interactor, surface sequencer.

The measured kernel is the mixed-precision batched kernel (FP32 stages,
field, transport; FP64 track state, path, error estimate, navigation) with
the launch picked by the autotuner; the plain FP64 kernel is timed on the
same batch for the speed-up, and the two are compared track by track
(surface sequences, bound parameters).  ``--workload loopers`` runs the
batching workload instead: 2M tracks, pT 0.1-100 GeV, |eta| < 3, in batches
of 262144 replayed from HIP graphs, with p50/p99 per-batch latency.

Prints one JSON line with the metric, the roofline of the propagation
kernel and a CPU baseline (the FP64 oracle, OpenMP on the host cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "tracks/s (toy detector, solenoid map) mixed-prec vs fp64 at <=1e-4 rel error"


def alg_bytes(n_surfaces):
    """Algorithmic HBM bytes of one launch: per track the FP64 free vector
    (64 B) and upper-triangular covariance (168 B) in, bound vector (48 B),
    covariance (168 B), status and counters (16 B) out, and 4+4 B per
    recorded surface (index in the sequence, and its share of the count)."""
    import numpy as np

    return float(np.sum(464 + 8 * n_surfaces.astype(np.int64)))


def traffic_bytes(headline: bool):
    """HBM bytes per launch of the propagation kernel measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by the copy-kernel
    calibration), as recorded in profiles/traffic.json for the headline
    workload; None if not recorded."""
    path = ROOT / "profiles" / "traffic.json"
    if not headline or not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def compare(mixed, ref):
    """Surface-sequence and bound-parameter agreement of two result sets
    (device tensors), reference convention: loc differences relative to the
    path length, angles relative to max(|x|, 1), q/p relative."""
    import torch

    ns_mm = (mixed.n_surfaces != ref.n_surfaces)
    ms = ref.surfaces.shape[0]
    k = torch.arange(ms, device=ref.surfaces.device)[:, None]
    valid = k < ref.n_surfaces.clamp(max=ms)[None, :]
    seq_mm = ((mixed.surfaces != ref.surfaces) & valid).any(dim=0) | ns_mm
    st_mm = mixed.status != ref.status
    path = ref.path_length.abs().clamp(min=1.0)
    bp, rp = mixed.bound_params, ref.bound_params
    d = (bp - rp).abs()
    scale = torch.stack([path, path, rp[2].abs().clamp(min=1.0), rp[3].abs().clamp(min=1.0),
                         torch.where(rp[4] != 0, rp[4].abs(), torch.ones_like(rp[4])),
                         rp[5].abs().clamp(min=1.0)])
    err = d / scale
    strict = d / rp.abs().clamp(min=1.0)
    strict[4] = d[4] / torch.where(rp[4] != 0, rp[4].abs(), torch.ones_like(rp[4]))
    per = err.max(dim=1).values.tolist()
    return {
        "tracks": int(rp.shape[1]),
        "surface_sequence_mismatches": int(seq_mm.sum().item()),
        "status_mismatches": int(st_mm.sum().item()),
        "max_rel_err": float(err.max().item()),
        "max_rel_err_per_param": [float(x) for x in per],
        "convention": "loc0/loc1 relative to path length, phi/theta to max(|x|,1), q/p relative "
                      "(propagator_test.hpp host/device convention)",
        "max_err_loc_over_max_abs_1mm": float(strict.max().item()),
        "p9999_err_loc_over_max_abs_1mm": float(torch.quantile(strict.max(dim=0).values.float(), 0.9999).item()),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tracks", type=int, default=None)
    ap.add_argument("--workload", choices=("toy_solenoid", "loopers"), default="toy_solenoid")
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2)
    # launch used with --no-autotune
    ap.add_argument("--sort", default="eta_phi")
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--waves", type=int, default=2)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "track ranges with no exchange, multi-GPU runs are not part of this build",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import detray_amd as da
    from detray_amd.schedule import CELL_KEYS, autotune, sort_order

    dev = torch.device("cuda:0")
    det = da.Detector.toy(4, 7)
    field = da.MagneticField.solenoid()
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    loopers = args.workload == "loopers"
    n = args.tracks or (2_000_000 if loopers else 1_000_000)
    seed = args.seed if args.seed is not None else (45 if loopers else 44)
    if loopers:
        fp_host = da.generate_tracks("toy", n, seed, pt_min=0.1, pt_max=100.0, eta_min=-3.0, eta_max=3.0)
    else:
        fp_host = da.generate_tracks("toy", n, seed)
    tracks = torch.from_numpy(fp_host).to(dev)
    print(f"[bench] {args.workload}: {n} tracks, toy 4brl/7edc, solenoid grid {field.name}",
          file=sys.stderr, flush=True)

    if args.no_autotune:
        launch = da.LaunchConfig("mixed", "grid", args.block, args.waves, 64, None, 1, "lane")
        key = args.sort
    else:
        t0 = time.perf_counter()
        launch, key, table = autotune(det, field, cfg, tracks, "mixed",
                                      sample=(262144 if loopers else n), tag=args.workload)
        print(f"[bench] autotune ({time.perf_counter() - t0:.1f} s): {launch.schedule} "
              f"nav={launch.navigation} block={launch.block_size} waves={launch.waves_per_simd} "
              f"idle={launch.min_idle} sort={key}", file=sys.stderr, flush=True)
    prop = da.Propagator(cfg, launch)
    stream = torch.cuda.current_stream(dev)

    batch = 262144 if loopers else n
    starts = list(range(0, n, batch))
    subs = [tracks[:, s:s + batch].contiguous() for s in starts]
    outs = [prop.allocate(t.shape[1], dev) for t in subs]

    graphs = None
    if loopers:
        # each batch (native sort + propagation) captured once in a HIP graph
        # and replayed; the graph holds kernels only (no allocation, no
        # memset)
        if key not in CELL_KEYS and key != "none":
            key = "cells"
        for i in range(len(subs)):
            launch.order = sort_order(subs[i], key)
            prop.propagate(det, field, subs[i], out=outs[i])
        torch.cuda.synchronize()
        graphs = []
        side = torch.cuda.Stream()
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            for i in range(len(subs)):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    launch.order = sort_order(subs[i], key)
                    prop.propagate(det, field, subs[i], out=outs[i])
                graphs.append(g)
        stream.wait_stream(side)
        torch.cuda.synchronize()

    def run_batch(i, ev=None):
        if ev is not None:
            ev[0].record(stream)
        if graphs is not None:
            if ev is not None:
                ev[1].record(stream)
            graphs[i].replay()
        else:
            launch.order = sort_order(subs[i], key)
            if ev is not None:
                ev[1].record(stream)
            prop.propagate(det, field, subs[i], out=outs[i])
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        for i in range(len(subs)):
            run_batch(i)
    torch.cuda.synchronize()

    evs = [[[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in subs] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        for i in range(len(subs)):
            run_batch(i, evs[s][i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n / (wall / args.steps)
    sort_ms = [e[0].elapsed_time(e[1]) for st in evs for e in st]
    kern_ms = [e[1].elapsed_time(e[2]) for st in evs for e in st]
    batch_ms = [e[0].elapsed_time(e[2]) for st in evs for e in st]

    n_sf = torch.cat([o.n_surfaces for o in outs]).cpu().numpy()
    status = torch.cat([o.status for o in outs]).cpu().numpy()
    n_steps = torch.cat([o.n_steps for o in outs]).cpu().numpy().astype(np.int64)
    bytes_per_step = alg_bytes(n_sf)
    kernel_s_per_step = float(np.sum(kern_ms)) / args.steps * 1e-3
    achieved = bytes_per_step / kernel_s_per_step / 1e9
    print(f"[bench] mixed: {ms_per_step:.2f} ms/step (sort {np.sum(sort_ms) / args.steps:.2f} ms, "
          f"kernel {np.sum(kern_ms) / args.steps:.2f} ms), {value:.3e} tracks/s, mean steps "
          f"{n_steps.mean():.1f}, mean N_sf {n_sf.mean():.2f}, status "
          f"{np.bincount(status, minlength=4).tolist()}", file=sys.stderr, flush=True)

    # the plain FP64 kernel on the same batches, for the speed-up and the
    # track-by-track comparison
    prop64 = da.Propagator(cfg)
    outs64 = [prop64.allocate(t.shape[1], dev) for t in subs]
    for i in range(len(subs)):
        prop64.propagate(det, field, subs[i], out=outs64[i])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        for i in range(len(subs)):
            prop64.propagate(det, field, subs[i], out=outs64[i])
    torch.cuda.synchronize()
    wall64 = (time.perf_counter() - t0) / args.fp64_steps
    value64 = n / wall64
    acc = [compare(outs[i], outs64[i]) for i in range(len(subs))]
    accuracy = {
        "vs": "plain FP64 kernel on the same tracks",
        "tracks": sum(a["tracks"] for a in acc),
        "surface_sequence_mismatches": sum(a["surface_sequence_mismatches"] for a in acc),
        "status_mismatches": sum(a["status_mismatches"] for a in acc),
        "max_rel_err": max(a["max_rel_err"] for a in acc),
        "max_rel_err_per_param": [max(a["max_rel_err_per_param"][k] for a in acc) for k in range(6)],
        "convention": acc[0]["convention"],
        "max_err_loc_over_max_abs_1mm": max(a["max_err_loc_over_max_abs_1mm"] for a in acc),
    }
    print(f"[bench] fp64 plain: {wall64 * 1e3:.2f} ms/step, {value64:.3e} tracks/s; speed-up "
          f"{value / value64:.2f}x; accuracy {accuracy}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        from tests import oracle_host

        m = min(args.cpu_sample, n)
        sample = np.ascontiguousarray(fp_host[:, :m])
        threads = os.cpu_count() or 1
        warm = max(1, m // 10)
        oracle_host.propagate(det, field, cfg, np.ascontiguousarray(sample[:, :warm]), threads=threads)
        times = []
        for _ in range(3):
            t1 = time.perf_counter()
            oracle_host.propagate(det, field, cfg, sample, threads=threads)
            times.append(time.perf_counter() - t1)
        cpu_rate = m / float(np.median(times))
        cpu = {"value": cpu_rate, "unit": "tracks/s", "cores": threads, "kind": "port",
               "sample": f"first {m} tracks of the same workload, FP64 oracle, OpenMP "
                         f"dynamic/1 over tracks, median of 3 runs"}
        print(f"[bench] cpu oracle {cpu_rate:.3e} tracks/s on {threads} threads", file=sys.stderr,
              flush=True)

    workload = ("toy detector 4brl+7edc, solenoid 5cm fp32 grid, 1M muons |eta|<2.5 "
                "pT 1-100 GeV log-uniform" if not loopers else
                "toy detector 4brl+7edc, solenoid 5cm fp32 grid, 2M muons |eta|<3 pT 0.1-100 GeV "
                "log-uniform, batches of 262144 replayed from HIP graphs")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {
            "workload": workload,
            "tracks": n,
            "kernel": f"propagate_{'coop_persistent' if launch.schedule == 'persistent' else ('grid' if launch.navigation == 'lane' else 'coop')}_kernel<mixed>",
            "launch": {"schedule": launch.schedule, "navigation": launch.navigation,
                       "block": launch.block_size, "waves_per_simd": launch.waves_per_simd,
                       "tracks_per_claim": launch.tracks_per_claim, "min_idle": launch.min_idle,
                       "sort": key},
            "parallelism": "1 GPU",
        },
        "fp64_plain": {"value": value64, "unit": "tracks/s", "kernel": "propagate_fp64_kernel"},
        "speedup_vs_fp64": value / value64,
        "accuracy": accuracy,
        "latency_ms": {"per": "batch" if loopers else "step",
                       "p50": float(np.percentile(batch_ms, 50)),
                       "p99": float(np.percentile(batch_ms, 99)),
                       "sort_mean": float(np.mean(sort_ms)),
                       "kernel_mean": float(np.mean(kern_ms))},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_bytes(not loopers),
            "alg_bytes_per_launch": bytes_per_step / len(subs),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
