"""The uneven batch and the served stream at their stated sizes.

Batch: 24 systems drawn as the uneven-batch workload states (N from
{14, 38, 54, 66, 114, 162}, r_s log-uniform in [1, 10], Jastrow b uniform in
[0.5, 2], 64-1024 walkers each), mixed precision, 200 measured sweeps through
a captured HIP graph replayed ten times.  Every system must come out exactly
as the same system swept alone by eager sweeps, and sampled walkers' last
local energies must match the FP64 oracle evaluated at their final positions
within 1e-6 x max(|E_L|, N x 1 mHa).

Stream: requests of N in {114, 162, 246, 294} with up to 4096 walkers served
three at a time by the continuous-batching server give the same results as
each request served alone, in FP64 and in mixed precision.
"""
import time

import numpy as np
import pytest

from helpers import oracle_system

pytestmark = pytest.mark.gpu


def batch_specs(seed=2024):
    rng = np.random.default_rng(seed)
    Ns = rng.choice([14, 38, 54, 66, 114, 162], size=24)
    rs = np.exp(rng.uniform(0.0, np.log(10.0), 24))
    bs = rng.uniform(0.5, 2.0, 24)
    Ws = rng.integers(64, 1025, 24)
    return [(int(n), float(r), float(b), int(w), 100 + i) for i, (n, r, b, w) in enumerate(zip(Ns, rs, bs, Ws))]


def test_uneven_batch_of_24_systems_equals_each_alone(vmc, oracle):
    import torch

    specs = batch_specs()
    assert {n for n, *_ in specs} == {14, 38, 54, 66, 114, 162}
    systems = [vmc.System.heg(n, rs, b, precision=vmc.PREC_MIXED, reinvert_interval=64) for n, rs, b, _, _ in specs]
    W = [w for *_, w, _ in specs]
    seeds = [s for *_, s in specs]
    step0 = [0.3 * rs for _, rs, _, _, _ in specs]
    stream = torch.cuda.Stream()
    b = vmc.Batch(systems, W, seeds, step0=step0, stream=stream.cuda_stream)
    b.init_positions()
    b.capture(20, measure=True, block_size=10)
    for _ in range(10):
        b.replay()
    torch.cuda.synchronize()
    got = []
    for i in range(len(specs)):
        m = b.member(i)
        got.append((m.positions(), m.stats(), m.steps()))
        assert m.sweeps_done() == 200
    b.close()
    print("batch of", len(specs), "systems,", sum(W), "walkers: 200 sweeps by graph replay", flush=True)
    for i, ((n, rs, bb, w, seed), s) in enumerate(zip(specs, systems)):
        e = vmc.Ensemble(s, w, seed, step0=step0[i])
        e.init_positions()
        for _ in range(10):
            e.sweep(20, measure=True, block_size=10)
        ref = (e.positions(), e.stats(), e.steps())
        e.close()
        for a, r in zip(got[i], ref):
            np.testing.assert_array_equal(a, r, err_msg=f"system {i} (N={n}, W={w})")
        st = got[i][1]
        assert np.all(st[:, vmc.STAT["NMEAS"]] == 200)
        assert np.all(st[:, vmc.STAT["BLK_N"]] == 20)
        # Two sampled walkers against the FP64 oracle at their positions.
        osys = oracle_system(oracle, s)
        for wk in (0, w - 1):
            e_or, _, _ = oracle.Walker(osys, got[i][0][wk]).local_energy()
            e_dev = st[wk, vmc.STAT["LAST_E"]]
            assert abs(e_dev - e_or) <= 1e-6 * max(abs(e_or), n * 1e-3), (i, n, wk, e_dev, e_or)
        print(f"system {i}: N={n} r_s={rs:.2f} b={bb:.2f} W={w} equal alone; E_L of walkers 0, {w - 1} "
              f"within 1e-6 of the oracle", flush=True)


LARGE = [  # (N, r_s, b, walkers, warmup, measure, seed)
    (114, 5.0, 0.8, 4096, 2, 6, 31),
    (162, 2.0, 1.2, 2048, 2, 6, 32),
    (246, 10.0, 1.5, 1024, 2, 4, 33),
    (294, 2.0, 1.0, 512, 2, 4, 34),
    (294, 15.0, 0.7, 4096, 1, 3, 35),
]


def serve(vmc, specs, precision, max_active, gaps=None):
    srv = vmc.Server(64 << 30, max_active=max_active, sweeps_per_turn=2)
    ids = []
    for k, (N, rs, b, W, warm, meas, seed) in enumerate(specs):
        s = vmc.System.heg(N, rs, b, precision=precision, reinvert_interval=64 if precision else 0)
        ids.append(srv.submit(s, W, warm, meas, seed, step0=0.3 * rs, block_size=2))
        if gaps is not None:
            time.sleep(gaps[k])
    out = {}
    while len(out) < len(specs):
        for r in srv.poll(wait=True):
            out[r["id"]] = r
    srv.close()
    return [out[i] for i in ids]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_server_large_requests_same_alone_and_in_the_mix(vmc, precision):
    alone = [serve(vmc, [spec], precision, 1)[0] for spec in LARGE]
    gaps = np.random.default_rng(9).exponential(0.01, len(LARGE))
    mix = serve(vmc, LARGE, precision, 3, gaps)
    for spec, a, m in zip(LARGE, alone, mix):
        assert a["status"] == 0 and m["status"] == 0, spec
        for k in ("mean_energy", "standard_error", "acceptance_rate", "moves"):
            assert a[k] == m[k], (spec, k, a[k], m[k])
        N, W, warm, meas = spec[0], spec[3], spec[4], spec[5]
        assert m["moves"] == N * W * (warm + meas)
        assert np.isfinite(m["mean_energy"])
    lat = sorted(r["latency_s"] for r in mix)
    print(f"precision {precision}: latencies {[round(x * 1e3, 1) for x in lat]} ms", flush=True)
