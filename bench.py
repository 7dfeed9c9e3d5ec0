"""Benchmark: tracks/s of detray's propagation hot path on one MI355X.

Workload (one "step" = one propagate() pass over the whole batch):
  toy detector (4 barrel layers, 7 endcap discs per side, beampipe, grids,
  homogeneous Si material), analytic finite-solenoid field (2 T centre,
  R = 1.2 m, half length 3 m) sampled on a 5 cm fp32 xyz grid with
  trilinear interpolation; 1M muons, eta ~ U[-2.5, 2.5], phi ~ U,
  pT log-uniform [1, 100] GeV, vertex sigma_xy = 10 um, sigma_z = 50 mm;
  search window 3x3, covariance transport, parameter_updater with
  pointwise material interactor, surface sequencer.

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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tracks", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=44)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "track ranges with no exchange, multi-GPU runs are not part of this build",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import detray_amd as da

    dev = torch.device("cuda:0")
    det = da.Detector.toy(4, 7)
    field = da.MagneticField.solenoid()
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    prop = da.Propagator(cfg)
    n = args.tracks
    fp_host = da.generate_tracks("toy", n, args.seed)
    tracks = torch.from_numpy(fp_host).to(dev)
    out = prop.allocate(n, dev)
    print(f"[bench] {n} tracks, toy 4brl/7edc, solenoid grid {field.name}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        prop.propagate(det, field, tracks, out=out)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        prop.propagate(det, field, tracks, out=out)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in ev]
    ms_per_step = 1e3 * wall / args.steps
    value = n / (wall / args.steps)

    n_sf = out.n_surfaces.cpu().numpy().astype(np.int64)
    status = out.status.cpu().numpy()
    n_steps = out.n_steps.cpu().numpy().astype(np.int64)
    # algorithmic bytes per track: free vector + upper-triangular covariance in,
    # bound vector + covariance + status/count out, 8 B per recorded surface
    alg_bytes = float(np.sum(464 + 8 * n_sf))
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    achieved = alg_bytes / avg_kernel_s / 1e9
    print(f"[bench] {ms_per_step:.2f} ms/step, kernel {np.mean(kernel_ms):.2f} ms, "
          f"{value:.3e} tracks/s, mean steps {n_steps.mean():.1f}, mean N_sf {n_sf.mean():.2f}, "
          f"status {np.bincount(status, minlength=4).tolist()}", file=sys.stderr, flush=True)

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

    line = {
        "metric": "tracks/s (toy detector, solenoid map) mixed-prec vs fp64 at <=1e-4 rel error",
        "value": value,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {
            "workload": "toy detector 4brl+7edc, solenoid 5cm fp32 grid, 1M muons "
                        "|eta|<2.5 pT 1-100 GeV log-uniform, FP64 kernel",
            "tracks": n,
            "kernel": "propagate_fp64_kernel",
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
