"""Benchmark: tracks/s of detray's propagation hot path on one MI355X.

Workloads (one "step" = one propagate() pass over the whole batch, the track
sort of the schedule included):

* ``toy_solenoid`` (default, the headline): toy detector (4 barrel layers, 7
  endcap discs per side, beampipe, grids, homogeneous Si material), analytic
  finite-solenoid field (2 T centre, R = 1.2 m, half length 3 m) sampled on
  a 5 cm fp32 xyz grid with trilinear interpolation; 1M muons, eta ~
  U[-2.5, 2.5], phi ~ U, pT log-uniform [1, 100] GeV, vertex sigma_xy =
  10 um, sigma_z = 50 mm; search window 3x3, covariance transport,
  parameter_updater with pointwise material interactor, surface sequencer.
  The measured kernel is the mixed-precision batched kernel (FP32 stages,
  field, transport; FP64 track state, path, error estimate, navigation)
  with the launch the autotuner picks; the plain FP64 kernel is timed on
  the same batch for the speed-up.
* ``toy_const``: the same detector in a constant 2 T field, 100k muons
  (seed 43); the plain FP64 GPU kernel against the CPU oracle.
* ``loopers``: 2M tracks, pT 0.1-100 GeV, |eta| < 3, in batches of 262144
  replayed from HIP graphs, with p50/p99 per-batch latency and the
  distribution of steps per track.
* ``wire_chamber``: the reference's wire-chamber benchmark case.
* ``adjoint``: 262144 tracks from the beam pipe (pT 0.5-50 GeV, |eta| <
  2.5), hits = true intersections of 1-sigma perturbed start parameters +
  smear (10 um, 50 um); one step is the recording forward pass and the
  adjoint backward pass (checkpoint every 32 steps, mixed precision) giving
  d chi2 / d(bound start parameters).

Accuracy is reported against the FP64 oracle on the CPU sample (surface
sequences, bound parameters, covariance) and against the plain FP64 GPU
kernel on every track.  Prints one JSON line with the metric, the HBM and
VALU rooflines of the propagation kernel and the CPU baseline (the FP64
oracle, OpenMP over the host cores, five timed repetitions).
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
# MI355X dense vector peaks (FP64, FP32), TFLOP/s
VALU_PEAK_FP64 = 78.6
VALU_PEAK_FP32 = 157.3
METRIC = "tracks/s (toy detector, solenoid map) mixed-prec vs fp64 at <=1e-4 rel error"
CPU_REPS = 5


def alg_bytes(n_surfaces):
    """Algorithmic HBM bytes of one launch: per track the FP64 free vector
    (64 B) and upper-triangular covariance (168 B) in, bound vector (48 B),
    covariance (168 B), status and counters (16 B) out, and 4+4 B per
    recorded surface (index in the sequence, and its share of the count)."""
    import numpy as np

    return float(np.sum(464 + 8 * n_surfaces.astype(np.int64)))


def alg_flops(counts, field_grid: bool):
    """Algorithmic FP64-equivalent flops per track of the reference's
    algorithm, from the oracle's per-track counts of Runge-Kutta trials,
    steps, recorded surfaces and navigation intersections:
    F = N_trial (100 + 2 F_B) + N_step (10 + F_B + 50 + 545)
      + N_sf (486 + 547 + 150 + 60 + 60) + N_isect 60,
    F_B = 75 for the trilinear field lookup (0 for a constant field)."""
    import numpy as np

    fb = 75.0 if field_grid else 0.0
    f = (counts["n_trials"] * (100.0 + 2.0 * fb) + counts["n_steps"] * (10.0 + fb + 50.0 + 545.0)
         + counts["n_surfaces"] * (486.0 + 547.0 + 150.0 + 60.0 + 60.0) + counts["n_isect"] * 60.0)
    return np.asarray(f, dtype=np.float64)


def measured_counters(kernel: str):
    """rocprofv3 counters of the headline kernel recorded in
    profiles/counters.json (HBM bytes per launch from FETCH_SIZE /
    WRITE_SIZE with the copy-kernel calibration, VALU busy, lane
    utilisation, occupancy); None if not recorded for this kernel."""
    path = ROOT / "profiles" / "counters.json"
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    return data.get(kernel)


def host_threads() -> int:
    """Host threads of the CPU baseline: the process's CPU share (OpenMP's
    setting when given), at most 16 (the share of one GPU on the box)."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(n, 16))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(det, field, cfg, fp_sample):
    """The FP64 oracle on the sample: timed CPU_REPS times after a warm-up of
    10 % of the sample (OpenMP, dynamic schedule, chunk 1, as the
    reference's CPU benchmark); then one counted run whose outputs are the
    accuracy reference and whose per-track counts feed the flop model."""
    import numpy as np

    from tests import oracle_host

    m = fp_sample.shape[1]
    threads = host_threads()
    oracle_host.propagate(det, field, cfg, np.ascontiguousarray(fp_sample[:, :max(1, m // 10)]), threads=threads)
    times = []
    for _ in range(CPU_REPS):
        t1 = time.perf_counter()
        oracle_host.propagate(det, field, cfg, fp_sample, threads=threads)
        times.append(time.perf_counter() - t1)
    ref = oracle_host.propagate_counted(det, field, cfg, fp_sample, threads=threads)
    t = np.array(times)
    rate = m / float(np.median(t))
    cpu = {"value": rate, "unit": "tracks/s", "cores": threads, "kind": "port",
           "sample": f"first {m} tracks of the same workload, FP64 oracle (the CPU restatement of the "
                     f"reference's propagation), OpenMP dynamic/1 over tracks, median of {CPU_REPS} runs "
                     f"after a 10% warm-up",
           "runs_s": [float(x) for x in t], "spread": float((t.max() - t.min()) / np.median(t)),
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    print(f"[bench] cpu oracle {rate:.3e} tracks/s on {threads} threads ({cpu_model()}), runs "
          f"{[round(x, 3) for x in times]} s", file=sys.stderr, flush=True)
    return cpu, ref


def _host(res, lo=0, hi=None):
    """Result tensors of tracks [lo, hi) as numpy arrays."""
    import numpy as np

    sl = slice(lo, hi)
    return {
        "bound_params": res.bound_params[:, sl].cpu().numpy(),
        "bound_cov": res.bound_cov[:, sl].cpu().numpy(),
        "path_length": res.path_length[sl].cpu().numpy(),
        "surfaces": res.surfaces[:, sl].cpu().numpy().view(np.uint32),
        "n_surfaces": res.n_surfaces[sl].cpu().numpy().view(np.uint32),
        "n_steps": res.n_steps[sl].cpu().numpy().view(np.uint32),
        "status": res.status[sl].cpu().numpy(),
    }


def accuracy_vs(got, ref, what: str):
    """Agreement of a result set with a reference on the same tracks
    (numpy dicts): surface sequences and statuses; bound-parameter error in
    the reference's host/device convention (propagator_test.hpp: loc
    relative to the path length, angles to max(|x|, 1), q/p relative) and
    loc relative to max(|loc|, 1 mm); covariance error |dC_ij| /
    sqrt(C_ii C_jj), per parameter and over tracks (max, p99.99)."""
    import numpy as np

    n = ref["status"].shape[0]
    ms = min(got["surfaces"].shape[0], ref["surfaces"].shape[0])
    k = np.arange(ms)[:, None]
    valid = k < np.minimum(ref["n_surfaces"], ms)[None, :]
    seq_mm = (((got["surfaces"][:ms] != ref["surfaces"][:ms]) & valid).any(axis=0)
              | (got["n_surfaces"] != ref["n_surfaces"]))
    bp, rp = got["bound_params"], ref["bound_params"]
    d = np.abs(bp - rp)
    path = np.maximum(np.abs(ref["path_length"]), 1.0)
    qs = np.where(rp[4] != 0, np.abs(rp[4]), 1.0)
    scale = np.stack([path, path, np.maximum(np.abs(rp[2]), 1.0), np.maximum(np.abs(rp[3]), 1.0), qs,
                      np.maximum(np.abs(rp[5]), 1.0)])
    err = d / scale
    strict = d / np.maximum(np.abs(rp), 1.0)
    strict[4] = d[4] / qs
    has = ref["n_surfaces"] > 0
    iu = np.triu_indices(6)
    diag = [k for k, (r, c) in enumerate(zip(*iu)) if r == c]
    dg = np.sqrt(np.abs(ref["bound_cov"][diag]))
    cov_norm = (dg[iu[0]] * dg[iu[1]]) + 1e-300
    cov_err = np.abs(got["bound_cov"] - ref["bound_cov"]) / cov_norm
    cov_track = cov_err[:, has].max(axis=0) if has.any() else np.zeros(1)
    return {
        "vs": what,
        "tracks": int(n),
        "surface_sequence_mismatches": int(seq_mm.sum()),
        "status_mismatches": int((got["status"] != ref["status"]).sum()),
        "max_rel_err": float(err[:, has].max(initial=0.0)),
        "max_rel_err_per_param": [float(x) for x in err[:, has].max(axis=1, initial=0.0)],
        "p9999_rel_err": float(np.quantile(err[:, has].max(axis=0), 0.9999)) if has.any() else 0.0,
        "convention": "loc0/loc1 relative to path length, phi/theta to max(|x|,1), q/p relative "
                      "(propagator_test.hpp host/device convention)",
        "max_err_loc_over_max_abs_1mm": float(strict[:, has].max(initial=0.0)),
        "cov_max_rel_err": float(cov_track.max()),
        "cov_p9999_rel_err": float(np.quantile(cov_track, 0.9999)),
        "cov_convention": "|dC_ij| / sqrt(C_ii C_jj), tracks with a recorded surface",
    }


def _merge_accuracy(parts):
    import numpy as np

    out = dict(parts[0])
    out["tracks"] = sum(p["tracks"] for p in parts)
    for k in ("surface_sequence_mismatches", "status_mismatches"):
        out[k] = sum(p[k] for p in parts)
    for k in ("max_rel_err", "max_err_loc_over_max_abs_1mm", "cov_max_rel_err", "p9999_rel_err",
              "cov_p9999_rel_err"):
        out[k] = max(p[k] for p in parts)
    out["max_rel_err_per_param"] = [max(p["max_rel_err_per_param"][i] for p in parts) for i in range(6)]
    return out


def _calibration_copy() -> None:
    import torch

    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("[bench] calibration copy: 1 GiB device to device", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tracks", type=int, default=None)
    ap.add_argument("--workload", choices=("toy_solenoid", "toy_const", "loopers", "adjoint", "wire_chamber",
                                               "toy_maps"),
                    default="toy_solenoid")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="tracks of the CPU oracle leg (default: the whole batch)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2)
    # launch used with --no-autotune
    ap.add_argument("--sort", default="cells_phi")
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--waves", type=int, default=2)
    ap.add_argument("--nav", default="gather")
    ap.add_argument("--xcd", type=int, default=64)
    # a 1 GiB device-to-device copy after the run: the kernel of known size
    # that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters
    ap.add_argument("--calibration-copy", action="store_true")
    args = ap.parse_args()
    if args.calibration_copy:
        import atexit

        atexit.register(_calibration_copy)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "track ranges with no exchange, multi-GPU runs are not part of this build",
              file=sys.stderr)
        return 2

    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "wire_chamber":
        return bench_wire_chamber(args)

    import numpy as np
    import torch

    import detray_amd as da
    from detray_amd.schedule import CELL_KEYS, autotune, sort_order

    dev = torch.device("cuda:0")
    maps = args.workload == "toy_maps"
    det = da.Detector.toy(4, 7, use_material="maps" if maps else True)
    loopers = args.workload == "loopers"
    const = args.workload == "toy_const"
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA) if const else da.MagneticField.solenoid()
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    n = args.tracks or (2_000_000 if loopers else (100_000 if const else 1_000_000))
    seed = args.seed if args.seed is not None else (45 if loopers else (43 if const else 44))
    if loopers:
        fp_host = da.generate_tracks("toy", n, seed, pt_min=0.1, pt_max=100.0, eta_min=-3.0, eta_max=3.0)
    else:
        fp_host = da.generate_tracks("toy", n, seed)
    tracks = torch.from_numpy(fp_host).to(dev)
    print(f"[bench] {args.workload}: {n} tracks, toy 4brl/7edc, field {field.name}", file=sys.stderr, flush=True)

    if args.no_autotune:
        launch = da.LaunchConfig("mixed", "grid", args.block, args.waves, 64, None, 1, args.nav, args.xcd)
        key = args.sort
    else:
        t0 = time.perf_counter()
        launch, key, table = autotune(det, field, cfg, tracks, "mixed",
                                      sample=(262144 if loopers else n), tag=args.workload)
        print(f"[bench] autotune ({time.perf_counter() - t0:.1f} s): {launch.schedule} "
              f"nav={launch.navigation} block={launch.block_size} waves={launch.waves_per_simd} "
              f"claim={launch.tracks_per_claim} idle={launch.min_idle} xcd={launch.xcd_chunk} sort={key}",
              file=sys.stderr, flush=True)
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
    kernel_s_per_launch = float(np.mean(kern_ms)) * 1e-3
    print(f"[bench] mixed: {ms_per_step:.2f} ms/step (sort {np.sum(sort_ms) / args.steps:.2f} ms, "
          f"kernel {np.sum(kern_ms) / args.steps:.2f} ms), {value:.3e} tracks/s, mean steps "
          f"{n_steps.mean():.1f} (p50 {np.percentile(n_steps, 50):.0f}, p99 {np.percentile(n_steps, 99):.0f}, "
          f"max {n_steps.max()}), mean N_sf {n_sf.mean():.2f}, status {np.bincount(status, minlength=4).tolist()}",
          file=sys.stderr, flush=True)

    # the plain FP64 kernel on the same batches, for the speed-up and the
    # track-by-track comparison
    prop64 = da.Propagator(cfg)
    outs64 = [prop64.allocate(t.shape[1], dev) for t in subs]
    for i in range(len(subs)):
        prop64.propagate(det, field, subs[i], out=outs64[i])
    torch.cuda.synchronize()
    ev64 = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.fp64_steps * len(subs))]
    t0 = time.perf_counter()
    for s in range(args.fp64_steps):
        for i in range(len(subs)):
            ev64[s * len(subs) + i][0].record(stream)
            prop64.propagate(det, field, subs[i], out=outs64[i])
            ev64[s * len(subs) + i][1].record(stream)
    torch.cuda.synchronize()
    wall64 = (time.perf_counter() - t0) / args.fp64_steps
    value64 = n / wall64
    kern64_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev64]))
    # the measured kernel: the mixed batched kernel, or for the constant-field
    # workload the plain FP64 kernel
    measured_s = kern64_ms * 1e-3 if const else kernel_s_per_launch
    achieved = bytes_per_step / len(subs) / measured_s / 1e9
    acc64 = _merge_accuracy([accuracy_vs(_host(outs[i]), _host(outs64[i]), "plain FP64 GPU kernel, every track")
                             for i in range(len(subs))])
    print(f"[bench] fp64 plain: {wall64 * 1e3:.2f} ms/step, {value64:.3e} tracks/s; speed-up "
          f"{value / value64:.2f}x; mixed vs fp64 kernel: {acc64['surface_sequence_mismatches']} sequence "
          f"mismatches, max bound err {acc64['max_rel_err']:.2e}, max cov err {acc64['cov_max_rel_err']:.2e}",
          file=sys.stderr, flush=True)

    cpu = None
    acc_oracle = None
    acc64_oracle = None
    valu = None
    # the whole batch (about 1 s per run for the headline's 1M tracks on 16
    # host threads); the accuracy comparison then covers every track
    m = min(args.cpu_sample or n, batch)
    if not args.no_cpu:
        cpu, ref = cpu_leg(det, field, cfg, np.ascontiguousarray(fp_host[:, :m]))
        acc_oracle = accuracy_vs(_host(outs[0], 0, m), ref, f"FP64 oracle, first {m} tracks")
        acc64_oracle = accuracy_vs(_host(outs64[0], 0, m), ref, f"FP64 oracle, first {m} tracks")
        print(f"[bench] mixed vs oracle: {acc_oracle}", file=sys.stderr, flush=True)
        print(f"[bench] plain fp64 vs oracle: {acc64_oracle}", file=sys.stderr, flush=True)
        # VALU roofline: the flop model on the oracle's counts of the sample,
        # scaled to the batch, over the kernel's time per launch
        flops_track = float(alg_flops(ref, not const).mean())
        flops_launch = flops_track * batch
        tf = flops_launch / measured_s / 1e12
        tf64 = flops_launch / (kern64_ms * 1e-3) / 1e12
        tf_mixed = flops_launch / kernel_s_per_launch / 1e12
        meas = measured_counters("propagate_gather_kernel<mixed,2>") if not (const or loopers or maps) else None
        valu = {"bound": "valu", "achieved": tf, "peak": VALU_PEAK_FP64, "unit": "TFLOP/s (fp64-equivalent)",
                "frac": tf / VALU_PEAK_FP64, "frac_of_fp32_peak": tf / VALU_PEAK_FP32,
                "flops_per_track": flops_track,
                "counts_per_track": {k: float(np.mean(ref[k])) for k in ("n_trials", "n_steps", "n_surfaces",
                                                                          "n_isect")},
                "plain_fp64_kernel": {"achieved": tf64, "frac": tf64 / VALU_PEAK_FP64},
                "mixed_kernel": {"achieved": tf_mixed, "frac": tf_mixed / VALU_PEAK_FP64},
                "measured": meas}
        print(f"[bench] VALU roofline: {flops_track:.0f} flop/track, {tf:.2f} TFLOP/s = "
              f"{100 * tf / VALU_PEAK_FP64:.1f} % of the FP64 vector peak", file=sys.stderr, flush=True)

    if loopers:
        workload = ("toy detector 4brl+7edc, solenoid 5cm fp32 grid, 2M muons |eta|<3 pT 0.1-100 GeV "
                    "log-uniform, batches of 262144 replayed from HIP graphs")
    elif const:
        workload = "toy detector 4brl+7edc, constant Bz 2 T, 100k muons |eta|<2.5 pT 1-100 GeV log-uniform"
    elif maps:
        workload = ("toy detector 4brl+7edc with material maps on every portal and the beampipe "
                    "(use_material_maps), solenoid 5cm fp32 grid, 1M muons |eta|<2.5 pT 1-100 GeV log-uniform")
    else:
        workload = ("toy detector 4brl+7edc, solenoid 5cm fp32 grid, 1M muons |eta|<2.5 "
                    "pT 1-100 GeV log-uniform")
    kname = {"persistent": "propagate_coop_persistent_kernel", "wave": "propagate_coop_kernel",
             "lane": "propagate_grid_kernel", "gather": "propagate_gather_kernel"}
    kernel = kname["persistent"] if launch.schedule == "persistent" else kname[launch.navigation]
    traffic = None
    if not (const or loopers or maps):
        meas = measured_counters("propagate_gather_kernel<mixed,2>")
        if meas and launch.navigation == "gather" and launch.waves_per_simd == 2:
            traffic = meas.get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC if not const else "tracks/s (toy detector, constant 2 T) plain FP64 GPU kernel vs CPU",
        "value": value if not const else value64,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps if not const else args.fp64_steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step if not const else wall64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if not const else "f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {
            "workload": workload,
            "tracks": n,
            "kernel": f"{kernel}<mixed>" if not const else "propagate_fp64_kernel",
            "launch": {"schedule": launch.schedule, "navigation": launch.navigation,
                       "block": launch.block_size, "waves_per_simd": launch.waves_per_simd,
                       "tracks_per_claim": launch.tracks_per_claim, "min_idle": launch.min_idle,
                       "xcd_chunk": launch.xcd_chunk, "sort": key},
            "parallelism": "1 GPU",
        },
        "fp64_plain": {"value": value64, "unit": "tracks/s", "kernel": "propagate_fp64_kernel",
                       "kernel_ms": kern64_ms},
        "mixed": {"value": value, "unit": "tracks/s", "kernel": f"{kernel}<mixed>",
                  "kernel_ms": float(np.mean(kern_ms))},
        "speedup_vs_fp64": value / value64,
        "accuracy": acc_oracle if acc_oracle is not None else acc64,
        "accuracy_vs_fp64_kernel": acc64,
        "accuracy_fp64_kernel_vs_oracle": acc64_oracle,
        "steps_per_track": {"mean": float(n_steps.mean()), "p50": float(np.percentile(n_steps, 50)),
                            "p99": float(np.percentile(n_steps, 99)), "max": int(n_steps.max())},
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
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_per_step / len(subs),
        },
        "valu_roofline": valu,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_wire_chamber(args) -> int:
    """The reference's WIRE_CHAMBER_W_COV_TRANSPORT benchmark case
    (tests/benchmarks/cpu/propagation.cpp:69-73, 143-146): the wire chamber
    with half length 500 mm, constant 2 T field, 512 x 512 tracks from the
    origin, |eta| < 3, 1-100 GeV, covariance transport, no noise estimation.
    The measured kernel is the mixed-precision batched kernel (grid
    schedule, per-lane navigation); the plain FP64 kernel is timed beside
    it on the same batch."""
    import numpy as np
    import torch

    import detray_amd as da

    dev = torch.device("cuda:0")
    det = da.Detector.wire_chamber(half_z=500.0)
    field = da.MagneticField.constant(0.0, 0.0, 2.0 * da.TESLA)
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    n = args.tracks or 512 * 512
    seed = args.seed if args.seed is not None else 48
    fp_host = da.generate_tracks("toy", n, seed, eta_min=-3.0, eta_max=3.0, sigma_xy=0.0, sigma_z=0.0)
    tracks = torch.from_numpy(fp_host).to(dev)
    stream = torch.cuda.current_stream(dev)

    def timed(prop, steps):
        out = prop.allocate(n, dev)
        for _ in range(args.warmup):
            prop.propagate(det, field, tracks, out=out)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
        t0 = time.perf_counter()
        for s in range(steps):
            evs[s][0].record(stream)
            prop.propagate(det, field, tracks, out=out)
            evs[s][1].record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, [e[0].elapsed_time(e[1]) for e in evs], out

    launch = da.LaunchConfig("mixed", "grid", args.block, 2, 64, None, 1, "gather")
    wall, kern_ms, out = timed(da.Propagator(cfg, launch), args.steps)
    wall64, kern64, out64 = timed(da.Propagator(cfg), max(1, args.fp64_steps))
    n_sf = out.n_surfaces.cpu().numpy()
    status = out.status.cpu().numpy()
    same_seq = bool(torch.equal(out.surfaces, out64.surfaces) and torch.equal(out.n_surfaces, out64.n_surfaces))
    value = n / wall
    bytes_per_step = alg_bytes(n_sf)
    achieved = bytes_per_step / (float(np.mean(kern_ms)) * 1e-3) / 1e9
    print(f"[bench] wire chamber: mixed {wall * 1e3:.2f} ms/step, {value:.3e} tracks/s; plain FP64 "
          f"{wall64 * 1e3:.2f} ms/step; mean N_sf {n_sf.mean():.2f}, status "
          f"{np.bincount(status, minlength=4).tolist()}, identical sequences {same_seq}", file=sys.stderr,
          flush=True)
    cpu = None
    acc_oracle = None
    if not args.no_cpu:
        m = min(args.cpu_sample or 100_000, n)
        cpu, ref = cpu_leg(det, field, cfg, np.ascontiguousarray(fp_host[:, :m]))
        acc_oracle = accuracy_vs(_host(out, 0, m), ref, f"FP64 oracle, first {m} tracks")
        print(f"[bench] wire chamber mixed vs oracle: {acc_oracle}", file=sys.stderr, flush=True)
    line = {
        "metric": "tracks/s (wire chamber, constant 2 T, covariance transport)",
        "value": value,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stages/transport, f64 state",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": "wire chamber 10 layers half_z 500 mm drift cells, tungsten rods, constant "
                               "2 T, 262144 muons from the origin |eta|<3 pT 1-100 GeV log-uniform",
                   "tracks": n, "kernel": "propagate_gather_kernel (mixed, extended geometry)",
                   "block": args.block, "parallelism": "1 GPU"},
        "fp64_plain": {"ms_per_step": wall64 * 1e3, "tracks_per_s": n / wall64,
                       "speedup_of_mixed": wall64 / wall, "identical_sequences": same_seq},
        "accuracy": acc_oracle,
        "latency_ms": {"per": "step", "p50": float(np.percentile(kern_ms, 50)),
                       "p99": float(np.percentile(kern_ms, 99))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": bytes_per_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def adjoint_alg_bytes(n_steps, n_events, n_hits, every):
    """Algorithmic HBM bytes of one forward + backward pass: per track the
    bound start (48 B) and start surface (4 B) read by both passes, the hits
    (20 B each + 4 B count) read by both, chi2 (8 B), status (4 B) and the
    gradient (48 B) written; the tape written once and read once: 16 B per
    step (step sizes), 56 B per checkpoint, 204 B per event (free state
    before 80 B, bound 48 B, free state after 56 B, five words), 12 B of
    counters; the segment replay writes and reads 56 B per step."""
    import numpy as np

    n_ckpt = (n_steps + every - 1) // every
    tape = 16 * n_steps + 56 * n_ckpt + 204 * n_events + 12
    per = 2 * (52 + 20 * n_hits + 4) + 8 + 4 + 48 + 2 * tape + 2 * 56 * n_steps
    return float(np.sum(per.astype(np.int64)))


def bench_adjoint(args) -> int:
    import numpy as np
    import torch

    import detray_amd as da
    from detray_amd import adjoint as adj

    dev = torch.device("cuda:0")
    det = da.Detector.toy(4, 7)
    field = da.MagneticField.solenoid()
    cfg = da.PropagationConfig(search_window=(3, 3), estimate_scattering_noise=False)
    n = args.tracks or 262144
    seed = args.seed if args.seed is not None else 46
    b0, start = adj.beampipe_starts(n, seed=seed)
    truth = adj.perturbed(b0, seed=seed + 1)
    bound = torch.from_numpy(b0).to(dev)
    start_d = torch.from_numpy(start).to(dev)
    acfg = da.AdjointConfig(precision="mixed")
    prop = da.Chi2Propagator(cfg, acfg)
    prop64 = da.Chi2Propagator(cfg, da.AdjointConfig(precision="fp64"))
    rec = prop64.record_hits(det, field, torch.from_numpy(truth).to(dev), start_d, 32)
    hits = rec.smeared(generator=torch.Generator(device=dev).manual_seed(seed + 2))
    print(f"[bench] adjoint: {n} tracks from the beam pipe, mean hits {rec.n_hits.float().mean().item():.2f}",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream(dev)

    def timed(p, steps, warmup):
        for _ in range(warmup):
            f = p.forward(det, field, bound, start_d, hits)
            p.backward(det, field, bound, start_d, hits, f)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
        t0 = time.perf_counter()
        for s in range(steps):
            evs[s][0].record(stream)
            f = p.forward(det, field, bound, start_d, hits)
            evs[s][1].record(stream)
            g = p.backward(det, field, bound, start_d, hits, f)
            evs[s][2].record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
        bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        return wall, fwd, bwd, f, g

    wall, fwd_ms, bwd_ms, f, g = timed(prop, args.steps, args.warmup)
    value = n / wall
    wall64, fwd64, bwd64, f64, g64 = timed(prop64, args.fp64_steps, 1)
    status = f.status.cpu().numpy()
    print(f"[bench] adjoint mixed: {wall * 1e3:.2f} ms/step (forward {fwd_ms:.2f}, backward {bwd_ms:.2f}) "
          f"{value:.3e} tracks/s; fp64 {wall64 * 1e3:.2f} ms/step; status {np.bincount(status, minlength=4).tolist()}",
          file=sys.stderr, flush=True)

    # accuracy on a sample: exact adjoint (mixed) against central finite
    # differences of the FP64 forward chi2, h = 1e-6 relative (tracks whose
    # finite difference changes by more than 1e-4 at h/2 are reported and
    # left out); transport adjoint against the Jacobian-chain gradient
    m = min(1024, n)
    sig = adj.start_sigma(b0[:, :m])
    sub_b = bound[:, :m].contiguous()
    sub_s = start_d[:m].contiguous()
    sub_h = da.Hits(hits.surface[:, :m].contiguous(), hits.loc[:, :, :m].contiguous(), hits.n_hits[:m].contiguous())

    def chi2_at(b):
        return prop64.forward(det, field, b, sub_s, sub_h).chi2

    def fd(rel):
        out = torch.zeros((6, m), dtype=torch.float64, device=dev)
        for i in range(5):
            hstep = rel * (sub_b[i].abs() if i == 4 else sub_b[i].abs().clamp(min=1.0))
            bp = sub_b.clone()
            bp[i] += hstep
            bm = sub_b.clone()
            bm[i] -= hstep
            out[i] = (chi2_at(bp) - chi2_at(bm)) / (2 * hstep)
        return out.cpu().numpy()

    def scaled(x, ref):
        d = (np.abs(x[:5] - ref[:5]) * sig[:5]).max(axis=0)
        return d / np.maximum((np.abs(ref[:5]) * sig[:5]).max(axis=0), 1e-300)

    fd1, fd2 = fd(1e-6), fd(5e-7)
    stable = scaled(fd2, fd1) <= 1e-4
    ok = (status[:m] == 0) & stable
    e_mixed = scaled(g[:, :m].cpu().numpy(), fd1)[ok]
    e_fp64 = scaled(g64[:, :m].cpu().numpy(), fd1)[ok]
    ft = prop.forward(det, field, sub_b, sub_s, sub_h, transport_gradient=True)
    gt = prop.backward(det, field, sub_b, sub_s, sub_h, ft, model="transport")
    e_chain = scaled(gt.cpu().numpy(), ft.grad_transport.cpu().numpy())[ft.status.cpu().numpy() == 0]
    e_chain_fd = scaled(ft.grad_transport.cpu().numpy(), fd1)[ok]
    all_ok = (status == 0) & (f64.status.cpu().numpy() == 0)
    sig_all = adj.start_sigma(b0)
    d_mf = (np.abs(g.cpu().numpy()[:5] - g64.cpu().numpy()[:5]) * sig_all[:5]).max(axis=0)
    e_mf = (d_mf / np.maximum((np.abs(g64.cpu().numpy()[:5]) * sig_all[:5]).max(axis=0), 1e-300))[all_ok]
    accuracy = {
        "metric": "max_i |g_i - r_i| sigma_i / max_i |r_i| sigma_i (sigma: 10 um, 10 um, 1 mrad, 1 mrad, 1% q/p)",
        "fd_sample_tracks": int(m),
        "fd_unstable_tracks": int((~stable).sum()),
        "exact_mixed_vs_fd": {"max": float(e_mixed.max()), "median": float(np.median(e_mixed))},
        "exact_fp64_vs_fd": {"max": float(e_fp64.max()), "median": float(np.median(e_fp64))},
        "transport_vs_jacobian_chain": {"max": float(e_chain.max()), "median": float(np.median(e_chain))},
        "jacobian_chain_vs_fd": {"max": float(e_chain_fd.max()), "median": float(np.median(e_chain_fd))},
        "mixed_vs_fp64_adjoint_all_tracks": {"max": float(e_mf.max()), "median": float(np.median(e_mf)),
                                             "p9999": float(np.quantile(e_mf, 0.9999))},
        "tape_status": np.bincount(status, minlength=4).tolist(),
    }
    print(f"[bench] adjoint accuracy {accuracy}", file=sys.stderr, flush=True)

    # algorithmic bytes from the tape counters
    st_steps, st_events = prop.tape_counters(f, n)
    n_steps = st_steps.cpu().numpy().astype(np.int64)
    n_events = st_events.cpu().numpy().astype(np.int64)
    n_hits = hits.n_hits.cpu().numpy().astype(np.int64)
    alg = adjoint_alg_bytes(n_steps, n_events, n_hits, acfg.checkpoint_every)
    achieved = alg / ((fwd_ms + bwd_ms) * 1e-3) / 1e9
    print(f"[bench] adjoint mean steps {n_steps.mean():.1f}, events {n_events.mean():.2f}, "
          f"{alg / n:.0f} B/track", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        from tests import oracle_host

        mc = min(args.cpu_sample or 100_000, n)
        threads = host_threads()
        bs = np.ascontiguousarray(b0[:, :mc])
        ss = np.ascontiguousarray(start[:mc].astype(np.uint32))
        oracle_host.propagate_hits(det, field, cfg, bs[:, : max(1, mc // 10)], ss[: max(1, mc // 10)], 32, threads)
        times = []
        for _ in range(CPU_REPS):
            t1 = time.perf_counter()
            oracle_host.propagate_hits(det, field, cfg, bs, ss, 32, threads)
            times.append(time.perf_counter() - t1)
        t = np.array(times)
        cpu = {"value": mc / float(np.median(t)), "unit": "tracks/s", "cores": threads, "kind": "port",
               "sample": f"first {mc} tracks: FP64 oracle propagation with covariance (Jacobian) transport "
                         f"and the sensitive-surface parameters, the reference's way to the chain "
                         f"gradient; OpenMP dynamic/1, median of {CPU_REPS} runs",
               "runs_s": [float(x) for x in t], "spread": float((t.max() - t.min()) / np.median(t)),
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}

    line = {
        "metric": "tracks/s of d chi2 / d(initial bound parameters) (forward + adjoint), toy detector, "
                  "solenoid map",
        "value": value,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded numpy generator; hits from the FP64 forward pass on perturbed parameters)",
        "config": {
            "workload": f"toy detector 4brl+7edc, solenoid 5cm fp32 grid, {n} muons from the beam pipe, "
                        "|eta|<2.5, pT 0.5-50 GeV log-uniform, hits smeared 10/50 um, checkpoint every 32 steps",
            "tracks": n,
            "kernels": "adjoint_record_kernel<mixed> + adjoint_backward_kernel<mixed>",
            "parallelism": "1 GPU",
        },
        "forward_ms": fwd_ms,
        "backward_ms": bwd_ms,
        "fp64": {"value": n / wall64, "unit": "tracks/s", "forward_ms": fwd64, "backward_ms": bwd64},
        "speedup_vs_fp64": wall64 / wall,
        "accuracy": accuracy,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "alg_bytes_per_launch": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
