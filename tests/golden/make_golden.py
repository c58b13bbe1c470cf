#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz from the reference build.

The reference is compiled from its own sources by `make -C oracle ref` into
oracle/_ref/libvmc_ref.so (oracle/ref_shim.cpp drives its classes).  This
script feeds that build seeded inputs and stores inputs and outputs only:
Slater (log|det|, D, inverse, ratio and Sherman-Morrison of one move,
derivatives), Jastrow (value, derivatives, delta, incremental derivatives),
EnergyTracker (G count, V_real, V_recip, constants, S_G, incremental update),
WaveFunction + EnergyTracker local energy, and short Simulation::run
trajectories (per-step local energies and positions).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

CASES = [(1, 5.0), (3, 8.25), (7, 6.0), (19, 9.0), (27, 12.0)]
SIMS = [(7, 6.0, 700, 400, 0.35, 12345, 50), (19, 9.0, 380, 300, 0.6, 2048, 100)]


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libvmc_ref.so missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20251202)
    out = {}
    for n, L in CASES:
        for c in range(4):
            key = f"n{n}_c{c}"
            pos = rng.uniform(0.0, L, (n, 3))
            p = int(rng.integers(0, n))
            newpos = pos[p] + rng.normal(0.0, 0.3, 3)
            newpos -= L * np.floor(newpos / L)
            a, b = 0.25 + 0.1 * c, 0.8 + 0.2 * c
            s = O.ref_slater_eval(pos, L)
            ratio, inv_after, d_after, g_sl, l_sl = O.ref_slater_move(pos, L, p, newpos)
            jv, jg, jl, jd, jg2, jl2 = O.ref_jastrow(pos, L, a, b, p, newpos)
            en = O.ref_energy(pos, L, p, newpos)
            lpsi, eg, el, e_loc = O.ref_local_energy(pos, L, a, b)
            out.update({
                f"{key}_pos": pos, f"{key}_L": np.array(L), f"{key}_moved": np.array(p),
                f"{key}_newpos": newpos, f"{key}_ab": np.array([a, b]),
                f"{key}_logdet": np.array(s.logdet), f"{key}_D": s.D, f"{key}_inv": s.inv,
                f"{key}_ratio": np.array(ratio), f"{key}_inv_after": inv_after,
                f"{key}_slater_grad": g_sl, f"{key}_slater_lap": l_sl,
                f"{key}_jastrow": np.array(jv), f"{key}_jastrow_grad": jg,
                f"{key}_jastrow_lap": jl, f"{key}_jastrow_delta": np.array(jd),
                f"{key}_jastrow_grad_after": jg2, f"{key}_jastrow_lap_after": jl2,
                f"{key}_ng": np.array(en["ng"]), f"{key}_vreal": np.array(en["vreal"]),
                f"{key}_vrecip": np.array(en["vrecip"]), f"{key}_self": np.array(en["self"]),
                f"{key}_bg": np.array(en["bg"]), f"{key}_s_re": en["s_re"],
                f"{key}_s_im": en["s_im"], f"{key}_vreal_after": np.array(en["vreal_after"]),
                f"{key}_vrecip_after": np.array(en["vrecip_after"]),
                f"{key}_log_psi": np.array(lpsi), f"{key}_grad": eg, f"{key}_lap": el,
                f"{key}_e_local": np.array(e_loc),
            })
    for i, (n, L, wu, ms, st, seed, blk) in enumerate(SIMS):
        r = O.ref_simulation_run(n, L, wu, ms, st, seed, blk, want_pos=True)
        out.update({
            f"sim{i}_args": np.array([n, L, wu, ms, st, seed, blk], dtype=np.float64),
            f"sim{i}_mean": np.array(r["mean"]), f"sim{i}_se": np.array(r["se"]),
            f"sim{i}_acc": np.array(r["acc"]), f"sim{i}_frames_e": r["frames_e"],
            f"sim{i}_frames_mean": r["frames_mean"], f"sim{i}_frames_se": r["frames_se"],
            f"sim{i}_last_pos": r["frames_pos"][-3 * n:],
        })
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
