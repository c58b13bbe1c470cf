"""Shared input generators and comparison helpers of the test suite."""
import numpy as np


def random_configs(rng, n_configs, N, L, step=0.4):
    """Uniform configurations in [0, L)^3 and one proposed move each: electron
    index and a Gaussian displacement wrapped back into the box."""
    pos = rng.uniform(0.0, L, size=(n_configs, N, 3))
    moved = rng.integers(0, N, size=n_configs).astype(np.int32)
    newpos = np.empty((n_configs, 3))
    for c in range(n_configs):
        q = pos[c, moved[c]] + rng.normal(0.0, step, 3)
        newpos[c] = q - L * np.floor(q / L)
    np.clip(newpos, 0.0, np.nextafter(L, 0.0), out=newpos)
    return pos, moved, newpos


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def oracle_frozen(O, sysd, pos, moved, newpos):
    """The oracle's view of what vmc_frozen_eval returns, per configuration."""
    out = []
    for c in range(pos.shape[0]):
        w = O.Walker(sysd, pos[c])
        ndet = 2 if sysd.n_dn > 0 else 1
        rec = dict(singular=w.singular)
        rec["logdet"] = [w.logdet(d) for d in range(ndet)]
        rec["inv_before"] = [w.inverse(d) for d in range(ndet)]
        rec["jastrow"] = w.jastrow()
        g, lp = w.derivatives()
        rec["grad"], rec["lap"] = g, lp
        rec["kinetic"] = w.kinetic()
        rec["vreal"], rec["vrecip"] = w.vreal(), w.vrecip()
        rec["s_k"] = np.stack(w.structure_factor(), 1)
        e, _, _ = w.local_energy()
        rec["e_local"] = e
        p = int(moved[c])
        rec["ratio"] = w.ratio(p, newpos[c])
        rec["delta_j"] = w.delta_j(p, newpos[c])
        w.accept(p, newpos[c])
        rec["vreal_inc"], rec["vrecip_inc"] = w.vreal(), w.vrecip()
        rec["inv_after"] = [w.inverse(d) for d in range(ndet)]
        rec["jastrow_after"] = w.jastrow()
        e1, t1, _ = w.local_energy()
        rec["e_local_after"], rec["kinetic_after"] = e1, t1
        out.append(rec)
    return out
