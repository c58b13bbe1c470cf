"""The FP64 oracle against vectors written by the reference build.

tests/golden/reference_vectors.npz holds inputs and outputs of the
reference's own classes (written by tests/golden/make_golden.py from
oracle/_ref/libvmc_ref.so).  The oracle must reproduce them to rounding: the
reference sums some loops in SIMD order, so agreement is ~1e-13 relative, and
Simulation::run trajectories (mt19937_64 + libstdc++ distributions) agree
move for move.  When the reference build itself is present the same checks
also run live on fresh random inputs.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "reference_vectors.npz")
CASES = [(1, 0), (1, 1), (1, 2), (1, 3), (3, 0), (3, 1), (3, 2), (3, 3), (7, 0), (7, 1), (7, 2),
         (7, 3), (19, 0), (19, 1), (19, 2), (19, 3), (27, 0), (27, 1), (27, 2), (27, 3)]


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def close(a, b, tol, floor=1.0):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= tol * np.maximum(np.abs(b), floor))


def oracle_case(O, pos, L, a, b, p, newpos):
    n = pos.shape[0]
    res = {}
    w = O.Walker(O.reference_system(n, L, a, b), pos)
    res["logdet"] = w.logdet()
    res["D"] = w.dmatrix()
    res["inv"] = w.inverse()
    res["jastrow"] = w.jastrow()
    res["vreal"], res["vrecip"] = w.vreal(), w.vrecip()
    re, im = w.structure_factor()
    res["s_re"], res["s_im"] = re, im
    res["ng"] = w.ng
    e, _, _ = w.local_energy()
    res["e_local"] = e
    g, lp = w.derivatives()
    res["grad"], res["lap"] = g, lp
    res["log_psi"] = w.logdet() + w.jastrow()
    res["ratio"] = w.ratio(p, newpos)
    res["jastrow_delta"] = w.delta_j(p, newpos)
    w0 = O.Walker(O.reference_system(n, L, 0.0, b), pos)
    gs, ls = w0.derivatives()
    res["slater_grad"], res["slater_lap"] = gs, ls
    res["jastrow_grad"], res["jastrow_lap"] = g - gs, lp - ls
    old = pos[p].copy()
    w.accept(p, newpos)
    res["inv_after"] = w.inverse()
    res["vreal_after"], res["vrecip_after"] = w.vreal(), w.vrecip()
    gj, lj = w.derivatives_after_step(True, p, old)
    return res


@pytest.mark.parametrize("n,c", CASES, ids=[f"n{n}_c{c}" for n, c in CASES])
def test_oracle_reproduces_reference_vectors(oracle, golden, n, c):
    k = f"n{n}_c{c}"
    pos, L = golden[f"{k}_pos"], float(golden[f"{k}_L"])
    a, b = golden[f"{k}_ab"]
    p, newpos = int(golden[f"{k}_moved"]), golden[f"{k}_newpos"]
    r = oracle_case(oracle, pos, L, a, b, p, newpos)
    assert r["ng"] == int(golden[f"{k}_ng"])
    tol = 1e-11
    for name in ("logdet", "jastrow", "vreal", "vrecip", "e_local", "log_psi", "ratio",
                 "jastrow_delta", "vreal_after", "vrecip_after"):
        assert close(r[name], golden[f"{k}_{name}"], tol), name
    for name in ("D", "inv", "inv_after", "s_re", "s_im", "grad", "lap", "slater_grad",
                 "slater_lap", "jastrow_grad", "jastrow_lap"):
        want = golden[f"{k}_{name}"]
        scale = max(float(np.abs(want).max()), 1.0)
        assert np.abs(r[name] - want).max() <= tol * scale, name


@pytest.mark.parametrize("i", [0, 1])
def test_oracle_simulation_reproduces_reference_trajectory(oracle, golden, i):
    n, L, wu, ms, st, seed, blk = golden[f"sim{i}_args"]
    n, wu, ms, seed, blk = int(n), int(wu), int(ms), int(seed), int(blk)
    out, fe, fm, fs, fa, fp = oracle.oracle_simulation_run(n, L, wu, ms, st, seed, blk,
                                                           want_pos=True)
    assert np.abs(fe - golden[f"sim{i}_frames_e"]).max() <= 1e-10
    assert np.abs(fp[-3 * n:] - golden[f"sim{i}_last_pos"]).max() <= 1e-12
    assert abs(out.mean_energy - float(golden[f"sim{i}_mean"])) <= 1e-10
    assert abs(out.acceptance_rate - float(golden[f"sim{i}_acc"])) == 0.0


ref_live = pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
    "libvmc_ref.so")), reason="reference build oracle/_ref absent")


@ref_live
@pytest.mark.parametrize("n", [2, 3, 7, 19, 27, 33])
def test_oracle_matches_live_reference(oracle, n):
    rng = np.random.default_rng(100 + n)
    for trial in range(6):
        L = float(rng.uniform(4.0, 15.0))
        a, b = float(rng.uniform(0.1, 0.6)), float(rng.uniform(0.5, 2.0))
        pos = rng.uniform(0.0, L, (n, 3))
        p = int(rng.integers(0, n))
        newpos = pos[p] + rng.normal(0.0, 0.4, 3)
        newpos -= L * np.floor(newpos / L)
        r = oracle_case(oracle, pos, L, a, b, p, newpos)
        s = oracle.ref_slater_eval(pos, L)
        ratio, inv_after, _, g_sl, l_sl = oracle.ref_slater_move(pos, L, p, newpos)
        jv, jg, jl, jd, _, _ = oracle.ref_jastrow(pos, L, a, b, p, newpos)
        en = oracle.ref_energy(pos, L, p, newpos)
        lpsi, eg, el, e_loc = oracle.ref_local_energy(pos, L, a, b)
        if not np.isfinite(s.logdet):
            continue
        assert close(r["logdet"], s.logdet, 1e-11)
        assert np.abs(r["inv"] - s.inv).max() <= 1e-10 * np.abs(s.inv).max()
        assert close(r["ratio"], ratio, 1e-10, 1e-3)
        assert np.abs(r["inv_after"] - inv_after).max() <= 1e-10 * np.abs(inv_after).max()
        assert close(r["jastrow"], jv, 1e-12) and close(r["jastrow_delta"], jd, 1e-12)
        for name in ("vreal", "vrecip", "vreal_after", "vrecip_after"):
            assert close(r[name], en[name], 1e-11), name
        assert close(r["e_local"], e_loc, 1e-10)
        assert np.abs(r["grad"] - eg).max() <= 1e-10 * max(np.abs(eg).max(), 1.0)


@ref_live
def test_oracle_trajectory_matches_live_reference(oracle):
    for n, L, wu, ms, st, seed, blk in [(7, 6.0, 210, 700, 0.4, 99, 35), (3, 5.0, 30, 500, 2.0, 7, 10)]:
        r = oracle.ref_simulation_run(n, L, wu, ms, st, seed, blk, want_pos=True)
        out, fe, fm, fs, fa, fp = oracle.oracle_simulation_run(n, L, wu, ms, st, seed, blk,
                                                               want_pos=True)
        assert np.abs(fe - r["frames_e"]).max() <= 1e-10
        assert np.abs(fp - r["frames_pos"]).max() <= 1e-12
        assert out.acceptance_rate == r["acc"]
        assert abs(out.mean_energy - r["mean"]) <= 1e-10 and abs(out.standard_error - r["se"]) <= 1e-10
