#!/usr/bin/env python3
"""Headline benchmark: electron-moves/s of the VMC sweep at N = 54
(27 up + 27 down), r_s = 2, b = 1, 4096 walkers on one MI355X, mixed
precision against the FP64 path, with the CPU baseline timed on the host's
cores in the same run.

A step is one sweep of every walker (each electron proposed once, in order)
followed by the FP64 re-inversion of the walkers due for it (mixed precision)
and the local energy of every walker.  Walker state is
resident in HBM before the timed region starts.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "variational-monte-carlo_amd"))

N_ELECTRONS, RS, JB, WALKERS = 54, 2.0, 1.0, 4096
SEED = 20251202
PROFILES = os.path.join(ROOT, "profiles")

# MI355X published dense peaks.
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_LDS_TBPS = 256 * 128 * 2.4e9 / 1e12  # 256 CUs x 128 B/clk x 2.4 GHz


def refine_flops_per_walker(n, ndet=2):
    """Algorithmic flops of the FP64 refinement of a walker's FP32 inverse in
    the mixed local energy (one Newton-Schulz step per determinant: the
    residual E = I - D Y and the correction Y E, 2 n^3 each)."""
    return ndet * 4 * n ** 3


def sweep_side_flops_per_move(n, nk, N, p):
    """Sweep-side work per proposed move of the reference algorithm on this
    workload (add/mul = 1, FMA = 2): orbital phases 6 n_k, ratio 2 n, new
    Jastrow pairs 40 (N-1), and on accept (rate p) the Sherman-Morrison update
    4 n^2."""
    return 6 * nk + 2 * n + 40 * (N - 1) + p * 4 * n * n


def measure_flops_per_walker(n, N, ng):
    """Algorithmic flops of the local-energy kernel for one walker (add/mul =
    1, FMA = 2; sqrt, reciprocal, exp and erfc counted as 1 each):
      unordered pairs N(N-1)/2 x 48   (shared: min image 3, r^2 5, r 1,
                                       1/(1+br) 2, u', u'' 6, erfc(ar)/r 9;
                                       each side: gradient 6, Laplacian 5)
      Slater terms N n x 14        (orbital phase 6, gradient 6, Laplacian 2)
      structure factor N N_G x 14 + 4 N_G   (e^{iG.r} from axis powers 12,
                                              accumulate 2; w |S|^2 4)"""
    return 24 * N * (N - 1) + 14 * N * n + 14 * N * ng + 4 * ng


def reinvert_flops_per_walker(n, ndet=2):
    """Algorithmic flops of one FP64 Newton-Schulz step per walker: E Y and
    Y (2I - E Y), 2 n^3 each, per determinant."""
    return ndet * 4 * n ** 3


def sweep_flops_per_move(n, N, p):
    """Algorithmic flops of one proposed move in the sweep kernel: orbital
    row 6 n, ratio 2 n, new Jastrow pair terms 15 (N-1); on accept (rate p)
    Sherman-Morrison 4 n^2 and the Jastrow row-sum update 15 (N-1)."""
    return 6 * n + 2 * n + 15 * (N - 1) + p * (4 * n * n + 15 * (N - 1))


def path_work_per_move(n, nk, N, ng, p):
    """Work per proposed move of the reference algorithm on this workload,
    independent of this implementation (add/mul = 1, FMA = 2): orbital phases
    6 n_k, ratio 2 n, new Jastrow pairs 40 (N-1); on accept (rate p) the
    Sherman-Morrison update 4 n^2, the incremental structure factor 26 N_G and
    the real-space energy update 60 (N-1); the local energy once per sweep,
    (16 n^2 + 30 N (N-1)) / N.  LDS bytes per move of an LDS-resident FP32
    design: ratio column 4 n, on accept the update 8 n^2 and S(G) 16 N_G, and
    three position passes 36 N.  Returns (flops, lds_bytes)."""
    flops = (6 * nk + 2 * n + 40 * (N - 1) + p * (4 * n * n + 26 * ng + 60 * (N - 1))
             + (16 * n * n + 30 * N * (N - 1)) / N)
    lds = 4 * n + p * (8 * n * n + 16 * ng) + 36 * N
    return flops, lds


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    try:
        with open(os.path.join(PROFILES, f"{kernel}_N54_W4096_traffic.md")) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


# Kernels of the headline step, their arithmetic and their algorithmic work.
KERNELS = {
    "mixed": {"sweep": ("k_sweep_rr", "f32", "sweep-side work of the reference per proposed move "
                        "(6 n_k + 2 n + 40 (N-1) + p 4 n^2) x 4096 x 54 moves"),
              "measure": ("k_measure_rr_mixed", "f64", "FP64 refinement of the FP32 inverse (8 n^3) and "
                          "the local energy (24 N(N-1) + 14 N n + 14 N N_G + 4 N_G) per walker x 4096"),
              "reinvert": ("k_reinv_rr_list", "f64", "Newton-Schulz step 8 n^3 per re-inverted walker")},
    "fp64": {"sweep": ("k_sweep_rr", "f64", "sweep-side work of the reference per proposed move x 4096 x 54"),
             "measure": ("k_measure_rr", "f64", "local energy per walker x 4096"),
             "reinvert": ("k_reinv_rr_list", "f64", "")},
}


def roofline_block(head, head_name):
    """Roofline entries of the step's kernels from their HIP-event times:
    the dominant one (largest time per step) is the roofline, the others
    follow as secondary entries.  Every kernel here is bound by VALU issue
    and latency (no MFMA, HBM traffic near its algorithmic bytes); `peak` is
    the published dense figure of the kernel's arithmetic (FP32 157.3 TFLOP/s
    counts packed FP32, which this build does not issue: non-packed FP32
    issues at the FP64 rate, 78.6 TFLOP/s, given as peak_nonpacked)."""
    ents = {}
    for kind in ("sweep", "measure", "reinvert"):
        if f"{kind}_ms_avg" not in head:
            continue
        name, arith, what = KERNELS[head_name][kind]
        peak = PEAK_FP32_TFLOPS if arith == "f32" else PEAK_FP64_TFLOPS
        achieved = head[f"{kind}_tflops"]
        per_step_ms = head[f"{kind}_ms_avg"] * head.get(f"{kind}_launches", 1) / max(head.get("sweep_launches", 1), 1)
        ents[kind] = {"kernel": name, "dtype": arith, "bound": "valu", "achieved": achieved, "peak": peak,
                      "unit": "TFLOP/s", "frac": achieved / peak, "peak_nonpacked": PEAK_FP64_TFLOPS,
                      "frac_nonpacked": achieved / PEAK_FP64_TFLOPS, "launch_ms_avg": head[f"{kind}_ms_avg"],
                      "ms_per_step": per_step_ms, "work": what}
    dom = max(ents, key=lambda k: ents[k]["ms_per_step"])
    line = dict(ents[dom])
    line["traffic"] = read_traffic(line["kernel"])
    line["secondary"] = {k: v for k, v in ents.items() if k != dom}
    acc_rate = head["acceptance"]
    f_move, lds_move = path_work_per_move(27, 14, 54, 257, acc_rate)
    line["path"] = {
        "what": "whole path at the measured electron-moves/s, work per move of the reference "
                "algorithm on this workload (n=27, n_k=14, N=54, N_G=257, acceptance p)",
        "flops_per_move": f_move, "achieved": f_move * head["moves_per_s"] / 1e12,
        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": f_move * head["moves_per_s"] / 1e12 / PEAK_FP64_TFLOPS,
        "lds_bytes_per_move": lds_move, "lds_achieved_TBps": lds_move * head["moves_per_s"] / 1e12,
        "lds_peak_TBps": PEAK_LDS_TBPS,
        "lds_frac": lds_move * head["moves_per_s"] / 1e12 / PEAK_LDS_TBPS}
    return line


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    s = sorted(xs)
    return s[min(len(s) - 1, int(q * len(s)))]


def gpu_run(V, torch, precision, steps, warmup, reinvert, kernel):
    s = V.System.heg(N_ELECTRONS, RS, JB, precision=precision, reinvert_interval=reinvert)
    (nk0, _), ng, _ = V.system_sizes(s)
    stream = torch.cuda.Stream()
    ens = V.Ensemble(s, WALKERS, SEED, step0=0.3 * RS, stream=stream.cuda_stream)
    ens.set_kernel(kernel)
    ens.init_positions()
    # Equilibration with step adaptation (untimed), then the measured sweeps.
    ens.sweep(max(warmup, 1), measure=False, adapt=True)
    ens.reset_stats()
    # Timed region 1: the throughput, as a production run issues it: one
    # call of `steps` measured sweeps, no events between launches.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.sweep(steps, measure=True, adapt=False, block_size=10)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # Timed region 2: the same steps again with HIP events on the launch
    # stream around every step and every kernel (latency and kernel times).
    ens.timing()  # drop anything recorded so far
    ens.set_timing(True)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        ens.sweep(1, measure=True, adapt=False, block_size=10)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall_ev = time.perf_counter() - t1
    ens.set_timing(False)
    kt = ens.timing()
    step_ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    st = ens.stats()
    drift = ens.drift()
    acc = float(st[:, V.STAT["ACCEPTED"]].sum() / st[:, V.STAT["PROPOSED"]].sum())
    reinv = float(st[:, V.STAT["REINV"]].sum()) / 2.0  # both timed regions
    mean, se = V.blocking_summary(st)
    used = ens.kernel()
    ens.close()
    n = N_ELECTRONS // 2
    moves = WALKERS * N_ELECTRONS * steps
    out = dict(moves_per_s=moves / wall, wall_s=wall, moves_per_s_with_events=moves / wall_ev,
               acceptance=acc,
               e_per_electron=mean / N_ELECTRONS,
               se_per_electron=None if se is None else se / N_ELECTRONS,
               step_ms_p50=pct(step_ms, 0.5), step_ms_p99=pct(step_ms, 0.99),
               step_ms_avg=sum(step_ms) / len(step_ms), kernel=used,
               reinversions_per_walker_step=reinv / (WALKERS * steps),
               drift_max=float(drift.max()))
    if kt["measure"][1]:
        ms = kt["measure"][0] / kt["measure"][1]
        f = measure_flops_per_walker(n, N_ELECTRONS, ng)
        if precision == V.PREC_MIXED:
            f += refine_flops_per_walker(n)
        out.update(measure_ms_avg=ms, measure_launches=kt["measure"][1],
                   measure_tflops=f * WALKERS / (ms * 1e-3) / 1e12, measure_flops_per_walker=f)
    if kt["reinvert"][1]:
        ms = kt["reinvert"][0] / kt["reinvert"][1]
        f = reinvert_flops_per_walker(n) * out["reinversions_per_walker_step"]
        out.update(reinvert_ms_avg=ms, reinvert_launches=kt["reinvert"][1],
                   reinvert_tflops=f * WALKERS / (ms * 1e-3) / 1e12)
    if kt["sweep"][1]:
        sweep_ms = kt["sweep"][0] / kt["sweep"][1]
        f = sweep_side_flops_per_move(n, nk0, N_ELECTRONS, acc)
        out.update(sweep_ms_avg=sweep_ms, sweep_launches=kt["sweep"][1],
                   sweep_tflops=f * WALKERS * N_ELECTRONS / (sweep_ms * 1e-3) / 1e12,
                   sweep_flops_per_move=f)
    return out


def physics_run(V, torch, precision, seed, warm=500, meas=5000, block=100):
    """The configuration's own run: N=54, r_s=2, b=1, 4096 walkers, `warm`
    adaptive equilibration sweeps then `meas` measured sweeps (E_L every
    sweep, blocks of `block`); energy per electron with its blocking error
    (walker aggregation of the reference's main.cpp:79-104)."""
    s = V.System.heg(N_ELECTRONS, RS, JB, precision=precision, reinvert_interval=256 if precision else 0)
    stream = torch.cuda.Stream()
    ens = V.Ensemble(s, WALKERS, seed, step0=0.3 * RS, stream=stream.cuda_stream)
    ens.init_positions()
    t0 = time.perf_counter()
    ens.sweep(warm, adapt=True)
    ens.reset_stats()
    ens.sweep(meas, measure=True, block_size=block)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = ens.stats()
    mean, se = V.blocking_summary(st)
    acc = float(st[:, V.STAT["ACCEPTED"]].sum() / st[:, V.STAT["PROPOSED"]].sum())
    ens.close()
    return dict(seed=seed, sweeps=f"{warm}+{meas}", block=block, e_per_electron=mean / N_ELECTRONS,
                se_per_electron=None if se is None else se / N_ELECTRONS, acceptance=acc, wall_s=wall)


def cpu_baseline(target_s):
    """The oracle's FP64 sweep twin (same system, proposals and per-sweep E_L
    as the GPU path), one chain per host core, sized to about target_s
    seconds of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    s = O.baseline_system(N_ELECTRONS, RS, JB)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    rate1, _ = O.bench_threads(s, SEED, 0.6, 1, 20)
    sweeps = max(20, int(target_s * rate1 / N_ELECTRONS))
    rate, wall = O.bench_threads(s, SEED, 0.6, cores, sweeps)
    return dict(value=rate, unit="electron-moves/s", cores=cores, kind="port",
                sample=f"{cores} independent chains x {sweeps} sweeps of N=54 (r_s=2, b=1), "
                       f"E_L every sweep, FP64 oracle sweep twin, {wall:.1f} s wall",
                single_core=rate1)


def reference_cpu(target_s):
    """The reference's own Simulation::run, compiled from its sources into
    oracle/_ref: its single determinant at N=27 (the closest closed shell to
    one spin of the workload), r_s=2, with the reference's Ewald rules and
    E_L after every move as the reference does.  As src/main.cpp:50-68 runs
    its chains, one independent Simulation per host core at once (ctypes
    releases the interpreter lock for each), after a one-core run.  None when
    that library was not built."""
    import threading

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libvmc_ref.so")):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    n = 27
    L = O.box_length_from_rs(RS, n)
    t0 = time.perf_counter()
    O.ref_simulation_run(n, L, 10, 200, 0.6, 11, 10)
    dt = time.perf_counter() - t0
    steps = max(200, int(target_s / dt * 210))
    t0 = time.perf_counter()
    O.ref_simulation_run(n, L, 10, steps, 0.6, 11, 10)
    dt1 = time.perf_counter() - t0
    threads = [threading.Thread(target=O.ref_simulation_run, args=(n, L, 10, steps, 0.6, 100 + i, 10))
               for i in range(cores)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    return dict(value=cores * (steps + 10) / wall, unit="electron-moves/s", cores=cores,
                single_core=(steps + 10) / dt1,
                sample=f"reference Simulation::run, N=27 polarised, r_s=2, {cores} independent chains x "
                       f"{steps} measured moves at once, {wall:.1f} s wall (one chain alone: {dt1:.1f} s)")


SERVE_NS = (14, 38, 54, 66, 114, 162, 246, 294)


def serving_requests(n_requests, seed):
    """Requests of the served stream: N from {14,38,54,66,114,162,246,294},
    r_s log-uniform in [1,20], b uniform in [0.5,2], walkers 128-4096,
    sweeps 100-2000 (a quarter of them warmup), seed."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_requests):
        N = int(rng.choice(SERVE_NS))
        rs = float(np.exp(rng.uniform(0.0, np.log(20.0))))
        b = float(rng.uniform(0.5, 2.0))
        W = int(rng.integers(128, 4097))
        sw = int(rng.integers(100, 2001))
        out.append((N, rs, b, W, sw, i + 1))
    return out


def serving_capacity(V, precision, sweeps=8):
    """Electron-moves/s of one request alone per N (1024 walkers, E_L every
    sweep, default kernels): the service-rate model that sets the offered
    load of the served stream."""
    rate = {}
    for N in SERVE_NS:
        s = V.System.heg(N, 2.0, 1.0, precision=precision, reinvert_interval=256 if precision else 0)
        e = V.Ensemble(s, 1024, 3, step0=0.6)
        e.init_positions()
        e.sweep(2, measure=True)
        e.positions()  # synchronises the stream
        t0 = time.perf_counter()
        e.sweep(sweeps, measure=True)
        e.positions()
        rate[N] = 1024 * N * sweeps / (time.perf_counter() - t0)
        e.close()
    return rate


def serving_run(V, precision, n_requests=16, seed=5, load=0.8, sweep_scale=1.0, rate=None):
    """Served request stream (serving_requests; sweeps scaled by
    `sweep_scale`, walkers as drawn) through the continuous-batching server,
    Poisson arrivals at the rate that offers `load` of the device: load /
    mean over requests of (moves / the alone throughput at its N).  Returns
    the request latency percentiles, the aggregate electron-moves/s and the
    offered and achieved load."""
    import numpy as np

    reqs = [(N, rs, b, W, max(8, int(sw * sweep_scale)), sd) for N, rs, b, W, sw, sd in serving_requests(n_requests, seed)]
    rate = rate or serving_capacity(V, precision)
    service = [N * W * sw / rate[N] for N, _, _, W, sw, _ in reqs]
    rate_hz = load / float(np.mean(service))
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate_hz, n_requests)
    srv = V.Server(128 << 30, max_active=16, sweeps_per_turn=4)
    t0 = time.perf_counter()
    for (N, rs, b, W, sw, sd), gap in zip(reqs, gaps):
        s = V.System.heg(N, rs, b, precision=precision, reinvert_interval=256 if precision else 0)
        srv.submit(s, W, sw // 4, sw - sw // 4, sd, step0=0.3 * rs, block_size=max(1, (sw - sw // 4) // 20))
        time.sleep(float(gap))
    res = []
    while len(res) < len(reqs):
        res += srv.poll(wait=True)
    wall = time.perf_counter() - t0
    srv.close()
    lat = sorted(r["latency_s"] for r in res)
    return dict(requests=len(res), failed=sum(r["status"] != 0 for r in res),
                p50_ms=1e3 * pct(lat, 0.5), p99_ms=1e3 * pct(lat, 0.99),
                moves_per_s=sum(r["moves"] for r in res) / wall, wall_s=wall,
                offered_rate_hz=rate_hz, offered_load=load, sweep_scale=sweep_scale,
                service_model_s=float(np.sum(service)),
                busy_fraction=float(np.sum(service)) / wall,
                alone_moves_per_s_1024_walkers={str(k): v for k, v in rate.items()})


def batch_run(V, torch, sweeps=20, seed=2024):
    """Uneven batch: 24 systems (N from {14,38,54,66,114,162}, r_s
    log-uniform in [1,10], b uniform in [0.5,2], 64-1024 walkers), mixed
    precision, grouped by kernel family (one launch per kernel per group,
    groups on forked streams), 10 sweeps per replay of a captured HIP graph."""
    import numpy as np

    rng = np.random.default_rng(seed)
    Ns = rng.choice([14, 38, 54, 66, 114, 162], size=24)
    rs = np.exp(rng.uniform(0.0, np.log(10.0), 24))
    bs = rng.uniform(0.5, 2.0, 24)
    Ws = rng.integers(64, 1025, 24)
    sy = [V.System.heg(int(n), float(r), float(b), precision=V.PREC_MIXED, reinvert_interval=256)
          for n, r, b in zip(Ns, rs, bs)]
    stream = torch.cuda.Stream()
    b = V.Batch(sy, Ws, list(range(1, 25)), step0=[0.3 * r for r in rs], stream=stream.cuda_stream,
                walker_threads=128)
    b.init_positions()
    b.sweep(3, adapt=True)
    b.capture(10, measure=True)
    b.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(sweeps // 10):
        b.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    b.close()
    moves = float(sum(int(n) * int(w) for n, w in zip(Ns, Ws))) * (sweeps // 10) * 10
    return dict(systems=24, walkers=int(Ws.sum()), moves_per_s=moves / dt,
                ms_per_sweep=1e3 * dt / ((sweeps // 10) * 10))


def configs_run(V, torch, sweeps=10):
    """Per-sweep time (E_L every sweep) of the other configurations on the
    default kernels: N=114 (r_s=5, b=0.8, 2048 walkers) in mixed precision
    at K = 16, 64, 256 and in FP64, and N = 66, 162, 246, 294 (r_s=2, 1024
    walkers) mixed at K=64 and FP64.  Per-kernel times from HIP events; for
    mixed precision the inverse drift met before re-inversions."""
    import numpy as np

    cases = [(114, 5.0, 0.8, V.PREC_MIXED, K, 2048) for K in (16, 64, 256)]
    cases.append((114, 5.0, 0.8, V.PREC_FP64, 0, 2048))
    for n in (66, 162, 246, 294):
        cases += [(n, 2.0, 1.0, V.PREC_MIXED, 64, 1024), (n, 2.0, 1.0, V.PREC_FP64, 0, 1024)]
    out = {}
    for n, rs, b, prec, K, W in cases:
        s = V.System.heg(n, rs, b, precision=prec, reinvert_interval=K)
        e = V.Ensemble(s, W, 3, step0=0.3 * rs)
        e.set_walker_threads(V.default_walker_threads(s))
        e.init_positions()
        e.sweep(5, adapt=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.sweep(sweeps, measure=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / sweeps
        e.timing()
        e.set_timing(True)
        e.sweep(sweeps, measure=True)
        torch.cuda.synchronize()
        kt = {k: v[0] / v[1] * 1e3 for k, v in e.timing().items() if v[1]}
        key = f"N{n}_{'mixed_K' + str(K) if prec == V.PREC_MIXED else 'fp64'}"
        kname = {V.KERNEL_REGISTER: "register", V.KERNEL_DELAYED: "delayed"}.get(e.kernel(), "generic")
        out[key] = dict(walkers=W, ms_per_sweep=dt * 1e3, moves_per_s=W * n / dt, kernel=kname,
                        us_per_launch={k: round(v, 1) for k, v in kt.items()})
        if prec == V.PREC_MIXED:
            d = e.drift()
            out[key]["drift"] = {"max": float(d.max()), "median": float(np.median(d))}
            # A measured mixed sweep refreshes every inverse in FP64 for its
            # local energy, so K acts on the unmeasured (equilibration)
            # sweeps: their time and re-inversions per walker and sweep.
            st0 = e.stats()[:, V.STAT["REINV"]].sum()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e.sweep(sweeps)
            torch.cuda.synchronize()
            dtu = (time.perf_counter() - t0) / sweeps
            out[key]["unmeasured"] = dict(
                ms_per_sweep=dtu * 1e3, moves_per_s=W * n / dtu,
                reinversions_per_walker_sweep=float(e.stats()[:, V.STAT["REINV"]].sum() - st0) / (W * sweeps))
        e.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--reinvert", type=int, default=256,
                    help="accepted moves between FP64 re-inversions in the mixed path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-stream, batch and other-configuration runs")
    ap.add_argument("--configs-only", action="store_true",
                    help="run only the other-configuration timings (for profiling them)")
    ap.add_argument("--no-physics", action="store_true",
                    help="skip the 500+5000-sweep physics runs (for profiling the timed kernels)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain device copies of 1 GiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the walkers are independent chains and one MI355X holds "
              f"the whole 4096-walker ensemble; --gpus {a.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import torch
    import vmc_amd as V

    V.device_check()
    if a.configs_only:
        print(json.dumps({"configs": configs_run(V, torch)}))
        return
    if a.copy_calibration:
        # Two plain 1 GiB copies for the HBM byte counters: the runtime's
        # device-to-device copy and an elementwise kernel (per-lane vector
        # loads and stores, like this library's kernels).
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    res = {}
    log("timed runs")
    if not a.fp64_only:
        res["mixed"] = gpu_run(V, torch, V.PREC_MIXED, a.steps, a.warmup, a.reinvert,
                               V.KERNEL_AUTO)
    if not a.mixed_only:
        res["fp64"] = gpu_run(V, torch, V.PREC_FP64, a.steps, a.warmup, 0, V.KERNEL_AUTO)
        res["fp64_generic"] = gpu_run(V, torch, V.PREC_FP64, a.steps, a.warmup, 0,
                                      V.KERNEL_GENERIC)
    head_name = "mixed" if "mixed" in res else "fp64"
    head = res[head_name]
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
        "dtype": "f32/f64" if head_name == "mixed" else "f64",
        "data": "synthetic: Philox uniform starts per walker, adaptive equilibration sweeps",
        "config": {"workload": "HEG VMC sweep N=54 (27+27), r_s=2, b=1, 4096 walkers, "
                               "alpha=5/L, |n|^2<=25 (257 G), rcut=L/2, Gaussian moves, "
                               "E_L every sweep",
                   "precision": "mixed: FP32 moves (orbitals, ratios, Sherman-Morrison, Jastrow "
                                "change), FP64 local energy from an FP64-refined inverse, FP64 "
                                "accumulators" if head_name == "mixed" else "fp64",
                   "reinvert_interval": a.reinvert if head_name == "mixed" else 0,
                   "kernel": "register" if head["kernel"] == V.KERNEL_REGISTER else "generic"},
        "timed_as": ("value: one call of `steps` measured sweeps; value_with_kernel_events: one call per "
                     "step with HIP events around every kernel"),
        "value_with_kernel_events": head["moves_per_s_with_events"],
        "latency_ms": {"p50": head["step_ms_p50"], "p99": head["step_ms_p99"],
                       "what": "one step (sweep, re-inversion, local energy) of all 4096 walkers, "
                               "HIP events on the launch stream"},
    }
    if "measure_ms_avg" in head:
        line["roofline"] = roofline_block(head, head_name)
    for k in ("fp64", "fp64_generic"):
        if k in res and head_name != k:
            r = res[k]
            line[k] = {"value": r["moves_per_s"], "step_ms_p50": r["step_ms_p50"],
                       "step_ms_p99": r["step_ms_p99"],
                       "kernel": {V.KERNEL_REGISTER: "register", V.KERNEL_DELAYED: "delayed"}.get(r["kernel"],
                                                                                               "generic")}
            if "measure_ms_avg" in r:
                line[k].update(measure_ms_avg=r["measure_ms_avg"], sweep_ms_avg=r["sweep_ms_avg"],
                               measure_frac_fp64_peak=r["measure_tflops"] / PEAK_FP64_TFLOPS)
    if "mixed" in res and "fp64" in res:
        # The target ratio: against the fastest FP64 kernel (register rows).
        line["mixed_over_fp64"] = res["mixed"]["moves_per_s"] / res["fp64"]["moves_per_s"]
        line["mixed_over_fp64_generic"] = (res["mixed"]["moves_per_s"] /
                                           res["fp64_generic"]["moves_per_s"])
    if not a.no_serving:
        log("served streams")
        # The full-size stream in mixed precision and in FP64 (the delayed-
        # update kernel at 123 and 147 per spin), each at 80% of its own
        # path's capacity.
        sv = {"mixed": serving_run(V, V.PREC_MIXED)}
        log("served stream, mixed: done")
        if not a.mixed_only:
            sv["fp64"] = serving_run(V, V.PREC_FP64)
            log("served stream, fp64: done")
        sv["what"] = ("16 requests of N from {14,38,54,66,114,162,246,294}, r_s log-uniform in [1,20], "
                      "b uniform in [0.5,2], walkers 128-4096, sweeps 100-2000 (a quarter warmup, E_L every "
                      "measured sweep), Poisson arrivals at 80% offered load by the path's alone-throughput "
                      "model")
        line["serving"] = sv
        line["latency_ms"]["p50_solve"] = sv["mixed"]["p50_ms"]
        line["latency_ms"]["p99_solve"] = sv["mixed"]["p99_ms"]
        line["latency_ms"]["solve_what"] = ("request latency (submit to result) of the full-size served "
                                            "stream (16 requests, 80% offered load), mixed")
        log("uneven batch")
        line["batch"] = batch_run(V, torch)
        log("other configurations")
        line["configs"] = configs_run(V, torch)
    # Physics at the configuration's stated length: four runs per path, every
    # run with its own seed, pooled (mean of the four, standard error
    # sqrt(sum se^2) / 4).
    log("physics runs")
    phys = {}
    for k, prec, off in (("mixed", V.PREC_MIXED, 1), ("fp64", V.PREC_FP64, 2)):
        if k not in res or a.no_physics:
            continue
        runs = [physics_run(V, torch, prec, SEED + off + 2 * i) for i in range(4)]
        ses = [r["se_per_electron"] for r in runs]
        phys[k] = {"runs": runs, "e_per_electron": sum(r["e_per_electron"] for r in runs) / len(runs),
                   "se_per_electron": (sum(x * x for x in ses) ** 0.5 / len(runs)) if all(ses) else None}
    if len(phys) == 2 and phys["mixed"]["se_per_electron"] and phys["fp64"]["se_per_electron"]:
        d = phys["mixed"]["e_per_electron"] - phys["fp64"]["e_per_electron"]
        phys["difference_in_sigma"] = abs(d) / (phys["mixed"]["se_per_electron"] ** 2 +
                                                phys["fp64"]["se_per_electron"] ** 2) ** 0.5
    line["physics"] = phys
    line["timed_runs"] = {k: {"acceptance": v["acceptance"],
                              "reinversions_per_walker_step": v["reinversions_per_walker_step"],
                              "drift_max": v["drift_max"]} for k, v in res.items()}
    if not a.no_cpu_baseline:
        log("CPU baseline")
        line["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        ref = reference_cpu(min(a.cpu_seconds, 5.0))
        if ref is not None:
            line["cpu_reference"] = ref
    print(json.dumps(line))


if __name__ == "__main__":
    main()
