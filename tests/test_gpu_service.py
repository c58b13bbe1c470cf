"""Host services over the ensemble API on the GPU: the autotuner and the
continuous-batching request server."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_working_setting(vmc):
    s = vmc.System.heg(54, 2.0, 1.0, precision=vmc.PREC_MIXED)
    t = vmc.autotune(s, 256, drift_tol=1e-2, sweeps=4)
    print("tuning", t, flush=True)
    assert t["kernel"] in (vmc.KERNEL_GENERIC, vmc.KERNEL_REGISTER)
    assert t["reinvert_interval"] in (16, 64, 256) and t["moves_per_s"] > 0
    # The N = 54 system has a register-resident instantiation, which is faster.
    assert t["kernel"] == vmc.KERNEL_REGISTER
    s2 = vmc.System.heg(54, 2.0, 1.0, precision=vmc.PREC_MIXED, reinvert_interval=t["reinvert_interval"])
    e = vmc.Ensemble(s2, 64, 3)
    e.set_kernel(t["kernel"])
    e.set_rr_variant(t["rr_variant"])
    e.init_positions()
    e.sweep(3, measure=True)
    assert np.all(np.isfinite(e.stats()[:, vmc.STAT["LAST_E"]]))
    e.close()
    # N = 66 (33 per spin) has the multi-wavefront register kernel, which
    # should win too; N = 186 (93 per spin) has no register instantiation,
    # so the tuner picks among the generic kernel's placements and workgroups.
    t = vmc.autotune(vmc.System.heg(66, 2.0, 1.0, precision=vmc.PREC_MIXED), 64, drift_tol=1e-2, sweeps=2)
    assert t["kernel"] == vmc.KERNEL_REGISTER
    t = vmc.autotune(vmc.System.heg(186, 3.0, 1.0, precision=vmc.PREC_MIXED), 16, drift_tol=1e-2, sweeps=1)
    assert t["kernel"] == vmc.KERNEL_GENERIC and t["inv_placement"] in (1, 2)
    assert t["walker_threads"] in (64, 128, 256)


REQUESTS = [  # (N, r_s, b, walkers, warmup, measure, seed)
    (14, 1.0, 1.0, 128, 6, 20, 11),
    (38, 3.0, 1.5, 64, 4, 12, 12),
    (54, 2.0, 1.0, 256, 5, 16, 13),
    (66, 2.0, 0.8, 64, 3, 10, 14),
    (14, 5.0, 0.6, 96, 6, 14, 15),
    (54, 8.0, 1.9, 128, 4, 18, 16),
]


def submit(vmc, srv, spec, precision):
    N, rs, b, W, warm, meas, seed = spec
    s = vmc.System.heg(N, rs, b, precision=precision, reinvert_interval=64 if precision else 0)
    return srv.submit(s, W, warm, meas, seed, step0=0.3 * rs, block_size=2)


def collect(srv, n):
    out = {}
    while len(out) < n:
        for r in srv.poll(wait=True):
            out[r["id"]] = r
    return out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_server_results_same_alone_and_in_the_mix(vmc, precision):
    alone = []
    for spec in REQUESTS:
        srv = vmc.Server(1 << 34, max_active=1)
        i = submit(vmc, srv, spec, precision)
        alone.append(collect(srv, 1)[i])
        srv.close()
    # The mix: arrivals spread over time, at most three running at once, so
    # requests are admitted while others are on the device.
    rng = np.random.default_rng(7)
    srv = vmc.Server(1 << 34, max_active=3, sweeps_per_turn=2)
    ids = []
    for spec in REQUESTS:
        ids.append(submit(vmc, srv, spec, precision))
        time.sleep(float(rng.exponential(0.002)))
    mix = collect(srv, len(REQUESTS))
    srv.close()
    for a, i in zip(alone, ids):
        m = mix[i]
        assert a["status"] == 0 and m["status"] == 0
        for k in ("mean_energy", "standard_error", "acceptance_rate", "moves"):
            assert a[k] == m[k], (k, a[k], m[k])
    lat = sorted(r["latency_s"] for r in mix.values())
    print(f"latencies p50 {lat[len(lat) // 2] * 1e3:.1f} ms max {lat[-1] * 1e3:.1f} ms", flush=True)


def test_server_admission_by_footprint(vmc):
    s = vmc.System.heg(54, 2.0, 1.0, precision=vmc.PREC_MIXED, reinvert_interval=64)
    one = vmc.footprint(s, 256)
    srv = vmc.Server(int(one * 1.5), max_active=8)  # room for one such request at a time
    ids = [srv.submit(s, 256, 2, 6, seed) for seed in (1, 2, 3)]
    res = collect(srv, 3)
    assert all(res[i]["status"] == 0 for i in ids)
    # The later requests waited for the earlier ones' memory.
    q = [res[i]["queued_s"] for i in ids]
    assert q[2] > q[0]
    with pytest.raises(vmc.VmcError):
        srv.submit(s, 1024, 2, 6, 9)  # larger than the whole budget
    srv.close()


def test_server_reports_a_failed_sweep(vmc):
    """A request whose measured sweeps fail (block size 0 is rejected by the
    sweep) reports that error and the sweeps it did, not a result; a good
    request served beside it is unaffected."""
    s = vmc.System.heg(14, 1.0, 1.0)
    srv = vmc.Server(1 << 34, max_active=4)
    bad = srv.submit(s, 32, 3, 5, 21, block_size=0)
    good = srv.submit(s, 32, 3, 5, 22, block_size=1)
    res = collect(srv, 2)
    srv.close()
    assert res[bad]["status"] == vmc.ERR_INVALID
    assert res[bad]["moves"] == 32 * 14 * 3  # the warmup sweeps only
    assert res[good]["status"] == 0
    assert res[good]["moves"] == 32 * 14 * 8


def test_server_cancel(vmc):
    """A running request ends after its turn in flight and a waiting one at
    once; both report ERR_CANCELLED with the moves their sweeps made, and the
    request admitted after them completes as it does alone."""
    s = vmc.System.heg(54, 2.0, 1.0, precision=vmc.PREC_MIXED, reinvert_interval=64)
    small = vmc.System.heg(14, 1.0, 1.0)
    srv = vmc.Server(1 << 34, max_active=1)
    long_id = srv.submit(s, 256, 2, 200000, 31)   # minutes of sweeps if left alone
    waiting = srv.submit(s, 256, 2, 10, 32)
    after = srv.submit(small, 32, 3, 5, 33, block_size=1)
    srv.cancel(waiting)
    time.sleep(0.3)
    srv.cancel(long_id)
    res = collect(srv, 3)
    with pytest.raises(vmc.VmcError):
        srv.cancel(long_id)       # already reported
    with pytest.raises(vmc.VmcError):
        srv.cancel(after + 100)   # never submitted
    srv.close()
    print({k: (v["status"], v["moves"], round(v["latency_s"], 3)) for k, v in res.items()}, flush=True)
    assert res[waiting]["status"] == vmc.ERR_CANCELLED and res[waiting]["moves"] == 0
    r = res[long_id]
    assert r["status"] == vmc.ERR_CANCELLED
    assert 0 < r["moves"] < 256 * 54 * 200002 and r["moves"] % (256 * 54) == 0
    assert r["latency_s"] < 30.0
    assert res[after]["status"] == 0 and res[after]["moves"] == 32 * 14 * 8
    alone = vmc.Server(1 << 34, max_active=1)
    i = alone.submit(small, 32, 3, 5, 33, block_size=1)
    ref = collect(alone, 1)[i]
    alone.close()
    for k in ("mean_energy", "standard_error", "acceptance_rate"):
        assert ref[k] == res[after][k], k
