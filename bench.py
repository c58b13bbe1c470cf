#!/usr/bin/env python3
"""Headline benchmark: electron-moves/s of the VMC sweep at N = 54
(27 up + 27 down), r_s = 2, b = 1, 4096 walkers on one MI355X, mixed
precision against the plain FP64 path, with the CPU baseline timed on the
host's cores in the same run.

A step is one sweep of every walker (each electron proposed once) followed by
the local-energy measurement; walker state stays resident in HBM/LDS and the
inputs are on the device before the timed region starts.  Prints one JSON
line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "variational-monte-carlo_amd"))

N_ELECTRONS, RS, JB, WALKERS = 54, 2.0, 1.0, 4096
SEED = 20251202

# MI355X published peaks.
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def flops_per_move(n, nk, N, ng, p):
    """Algorithmic flops of one proposed move (add/mul = 1, FMA = 2,
    transcendentals not counted): orbital phases 6 n_k, ratio 2 n,
    Jastrow pairs 40 (N-1), on accept p [SM 4 n^2 + rho_k 26 N_G + real-space
    60 (N-1)], per-sweep local energy [16 n^2 + 30 N (N-1)] / N."""
    return (6 * nk + 2 * n + 40 * (N - 1) + p * (4 * n * n + 26 * ng + 60 * (N - 1)) +
            (16 * n * n + 30 * N * (N - 1)) / N)


def gpu_run(V, torch, precision, steps, warmup, sweeps_per_launch, reinvert):
    s = V.System.heg(N_ELECTRONS, RS, JB, precision=precision, reinvert_interval=reinvert)
    (nk0, _), ng, _ = V.system_sizes(s)
    stream = torch.cuda.Stream()
    ens = V.Ensemble(s, WALKERS, SEED, step0=0.3 * RS, stream=stream.cuda_stream)
    ens.init_positions()
    # Equilibration with step adaptation (untimed), then the measured sweeps.
    ens.sweep(max(warmup, 1), measure=False, adapt=True)
    ens.reset_stats()
    launches = math.ceil(steps / sweeps_per_launch)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(launches)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    for i in range(launches):
        k = min(sweeps_per_launch, steps - done)
        starts[i].record(stream)
        ens.sweep(k, measure=True, adapt=False, block_size=10)
        ends[i].record(stream)
        done += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    st = ens.stats()
    acc = float(st[:, V.STAT["ACCEPTED"]].sum() / st[:, V.STAT["PROPOSED"]].sum())
    mean, se = V.blocking_summary(st)
    moves = WALKERS * N_ELECTRONS * steps
    n = N_ELECTRONS // 2
    f_move = flops_per_move(n, nk0, N_ELECTRONS, ng, acc)
    # Per-launch algorithmic flops over the event-timed launch durations.
    per_launch = [(min(sweeps_per_launch, steps - i * sweeps_per_launch)) for i in range(launches)]
    tflops = [f_move * WALKERS * N_ELECTRONS * k / (ms * 1e-3) / 1e12
              for k, ms in zip(per_launch, launch_ms)]
    ens.close()
    srt = sorted(launch_ms)
    return dict(moves_per_s=moves / wall, wall_s=wall, acceptance=acc, e_per_electron=mean /
                N_ELECTRONS, se_per_electron=None if se is None else se / N_ELECTRONS,
                launch_ms_avg=sum(launch_ms) / len(launch_ms),
                launch_ms_p50=srt[len(srt) // 2], launch_ms_p99=srt[min(len(srt) - 1,
                                                                       int(0.99 * len(srt)))],
                tflops=sum(tflops) / len(tflops), flops_per_move=f_move, sweeps=steps,
                sweeps_per_launch=sweeps_per_launch)


def cpu_baseline(target_s):
    """The oracle's sweep twin, one chain per host core, sized to about
    target_s seconds of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    s = O.baseline_system(N_ELECTRONS, RS, JB)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    rate1, el = O.bench_threads(s, SEED, 0.6, 1, 20)
    sweeps = max(20, int(target_s * rate1 / N_ELECTRONS))
    rate, wall = O.bench_threads(s, SEED, 0.6, cores, sweeps)
    return dict(value=rate, unit="electron-moves/s", cores=cores, kind="port",
                sample=f"{cores} independent chains x {sweeps} sweeps of N=54 (r_s=2, b=1), "
                       f"E_L every sweep, FP64 oracle sweep twin, {wall:.1f} s wall",
                single_core=rate1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--sweeps-per-launch", type=int, default=10)
    ap.add_argument("--reinvert", type=int, default=64,
                    help="accepted moves between FP64 re-inversions in the mixed path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: walkers are independent chains and one MI355X runs the "
              f"whole 4096-walker ensemble; --gpus {a.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import torch
    import vmc_amd as V

    V.device_check()
    res = {}
    if not a.fp64_only:
        res["mixed"] = gpu_run(V, torch, V.PREC_MIXED, a.steps, a.warmup, a.sweeps_per_launch,
                               a.reinvert)
    if not a.mixed_only:
        res["fp64"] = gpu_run(V, torch, V.PREC_FP64, a.steps, a.warmup, a.sweeps_per_launch, 0)
    head = res.get("mixed", res.get("fp64"))
    head_name = "mixed" if "mixed" in res else "fp64"
    peak = PEAK_FP32_TFLOPS if head_name == "mixed" else PEAK_FP64_TFLOPS
    line = {
        "metric": "electron-moves/s at N=54, 4096 walkers, FP32 vs FP64; p50/p99 solve latency",
        "value": head["moves_per_s"],
        "unit": "electron-moves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * head["wall_s"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators and re-inversion)" if head_name == "mixed" else "f64",
        "data": "synthetic: Philox uniform starts, 50 adaptive equilibration sweeps",
        "config": {"workload": "HEG VMC sweep N=54 (27+27), r_s=2, b=1, 4096 walkers, "
                               "alpha=5/L, |n|^2<=25, rcut=L/2, Gaussian moves, E_L per sweep",
                   "precision": head_name, "sweeps_per_launch": a.sweeps_per_launch,
                   "reinvert_interval": a.reinvert if head_name == "mixed" else 0},
        "latency_ms": {"p50": head["launch_ms_p50"], "p99": head["launch_ms_p99"],
                       "what": f"one launch of {a.sweeps_per_launch} sweeps of all 4096 walkers"},
        "roofline": {"bound": "mfma", "engine": "valu (no MFMA in the sweep)",
                     "achieved": head["tflops"], "peak": peak, "unit": "TFLOP/s",
                     "frac": head["tflops"] / peak, "traffic": None,
                     "flops_per_move": head["flops_per_move"]},
    }
    if "mixed" in res and "fp64" in res:
        line["fp64"] = {"value": res["fp64"]["moves_per_s"],
                        "tflops": res["fp64"]["tflops"],
                        "frac_fp64_peak": res["fp64"]["tflops"] / PEAK_FP64_TFLOPS,
                        "launch_ms_p50": res["fp64"]["launch_ms_p50"],
                        "launch_ms_p99": res["fp64"]["launch_ms_p99"],
                        "e_per_electron": res["fp64"]["e_per_electron"],
                        "se_per_electron": res["fp64"]["se_per_electron"]}
        line["mixed_over_fp64"] = res["mixed"]["moves_per_s"] / res["fp64"]["moves_per_s"]
    line["physics"] = {k: {"e_per_electron": v["e_per_electron"],
                           "se_per_electron": v["se_per_electron"],
                           "acceptance": v["acceptance"]} for k, v in res.items()}
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
