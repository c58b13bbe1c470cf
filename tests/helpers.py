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


def oracle_system(O, s):
    """The oracle's orc_system for a vmc_amd.System."""
    return O.System(s.n_up, s.n_dn, s.box_length, s.a_same, s.a_anti, s.b, s.alpha,
                    s.recip_nmax2, s.real_cutoff, s.erfc_kind)


def sweep_order(n_up, n_dn):
    """Electron order of one sweep: 0..N-1, or with equal spins alternating
    up 0, down 0, up 1, ... (the device's sweep_visit)."""
    N = n_up + n_dn
    if n_up != n_dn:
        return list(range(N))
    return [(t >> 1) + (n_up if t & 1 else 0) for t in range(N)]


def oracle_given_sweep(O, sysd, pos, prop, uniforms, forced=None):
    """The oracle's sweep of each configuration with the given proposals
    [C][N][3]: per move the ratio, Jastrow change and decision
    u < R^2 exp(2 dJ) (or forced[c][p] when given), then the state after the
    sweep.  Returns per configuration a dict: trace [N][3] by electron, the
    rounding scale of each move's ratio (sum_m |phi_m| |invT[p][m]|), the
    inverse per determinant, positions, logdet, kinetic, vreal, vrecip,
    e_local, jastrow."""
    out = []
    order = sweep_order(sysd.n_up, sysd.n_dn)
    ndet = 2 if sysd.n_dn > 0 else 1
    n_det = [sysd.n_up, sysd.n_dn]
    shells = [O.shell_fill(n) if n else None for n in n_det]
    for c in range(pos.shape[0]):
        w = O.Walker(sysd, pos[c])
        N = pos.shape[1]
        tr = np.zeros((N, 3))
        scale = np.zeros(N)
        for p in order:
            q = prop[c, p]
            # The ratio's rounding scale: sum_m |phi_m(q)| |invT[p][m]|.
            d = 1 if (ndet == 2 and p >= sysd.n_up) else 0
            r = p - d * sysd.n_up
            nvec, orbk, orbt = shells[d]
            ph = 2.0 * np.pi * (nvec[orbk] @ np.asarray(q)) / sysd.L
            phi = np.where(orbt == 1, np.sin(ph), np.cos(ph))
            scale[p] = float(np.abs(phi) @ np.abs(w.inverse(d)[r]))
            R = w.ratio(p, q)
            dj = w.delta_j(p, q)
            if forced is not None:
                acc = bool(forced[c, p])
            else:
                acc = bool(uniforms[c, p] < R * R * np.exp(2.0 * dj))
            tr[p] = (R, dj, 1.0 if acc else 0.0)
            if acc:
                w.accept(p, q)
        e, t, _ = w.local_energy()
        out.append(dict(trace=tr, ratio_scale=scale, inverse=[w.inverse(d) for d in range(ndet)],
                        positions=w.positions(), logdet=[w.logdet(d) for d in range(ndet)],
                        kinetic=t, vreal=w.vreal(), vrecip=w.vrecip(), e_local=e,
                        jastrow=w.jastrow()))
    return out
