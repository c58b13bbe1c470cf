"""Pins of the oracle's extensions beyond the reference: the spin-resolved
Jastrow factor's derivatives, the benchmark configurations' Ewald rules
(alpha = 5/L, |n|^2 <= 25, spherical real-space cutoff L/2) and the
unpolarised Jastrow-free chain.  Each is checked against what mathematics
fixes: finite differences of the pinned values, the alpha-independence of a
converged Ewald sum (with the truncation's mean-field size), and the
closed-form Hartree-Fock energy of the plane-wave determinant."""
import numpy as np
import pytest

from hf_energy import box_length, ewald_truncation, hartree_fock_energy


def unpolarised(O, n_total, rs, a_same, a_anti, b, baseline=True):
    L = box_length(n_total, rs)
    if baseline:
        return O.System(n_total // 2, n_total // 2, L, a_same, a_anti, b, 5.0 / L, 25, 0.5 * L, 1)
    return O.System(n_total // 2, n_total // 2, L, a_same, a_anti, b, 8.0 / L, 0, 0.0, 1)


def ln_psi(O, sysd, pos):
    w = O.Walker(sysd, pos)
    return w.logdet(0) + w.logdet(1) + w.jastrow()


def spread_configs(rng, n_configs, N, L, dmin):
    """Uniform configurations with every minimum-image pair at least dmin
    apart (the finite differences stay away from the cusps)."""
    out = []
    while len(out) < n_configs:
        p = rng.uniform(0.0, L, (N, 3))
        d = p[:, None, :] - p[None, :, :]
        d -= L * np.rint(d / L)
        r = np.sqrt((d ** 2).sum(-1)) + np.eye(N) * L
        if r.min() > dmin:
            out.append(p)
    return out


def test_spin_resolved_derivatives_match_finite_differences(oracle):
    """grad and Laplacian of ln|psi| (two determinants and the Jastrow factor
    with a_same != a_anti) against central differences of ln|D_up| +
    ln|D_down| + J (values pinned to the reference by composition): a
    derivative that took the wrong spin pair's cusp misses by ~0.1."""
    O = oracle
    sysd = unpolarised(O, 14, 2.0, 0.25, 0.5, 0.8)
    rng = np.random.default_rng(71)
    for pos in spread_configs(rng, 4, 14, sysd.L, 0.6):
        w = O.Walker(sysd, pos)
        g, lap = w.derivatives()
        for i in range(14):
            fd_lap = 0.0
            f0 = ln_psi(O, sysd, pos)
            for c in range(3):
                h = 1e-5
                pp, pm = pos.copy(), pos.copy()
                pp[i, c] += h
                pm[i, c] -= h
                fd = (ln_psi(O, sysd, pp) - ln_psi(O, sysd, pm)) / (2 * h)
                assert abs(g[i, c] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, c, g[i, c], fd)
                h2 = 1e-4
                pp, pm = pos.copy(), pos.copy()
                pp[i, c] += h2
                pm[i, c] -= h2
                fd_lap += (ln_psi(O, sysd, pp) - 2 * f0 + ln_psi(O, sysd, pm)) / h2 ** 2
            assert abs(lap[i] - fd_lap) <= 5e-4 * max(1.0, abs(fd_lap)), (i, lap[i], fd_lap)


def test_spin_resolved_jastrow_uses_each_pairs_cusp(oracle):
    """J of the unpolarised system equals the sum over pairs of
    a r / (1 + b r), a = a_same for parallel and a_anti for antiparallel
    spins (a direct double loop here), and the derivative of J alone,
    ln|psi| of a Jastrow-only change, follows it."""
    O = oracle
    sysd = unpolarised(O, 14, 2.0, 0.25, 0.5, 0.8)
    L = sysd.L
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, L, (14, 3))
    J = 0.0
    for i in range(14):
        for j in range(i + 1, 14):
            d = pos[i] - pos[j]
            d -= L * np.rint(d / L)
            r = np.sqrt(d @ d)
            a = 0.25 if (i < 7) == (j < 7) else 0.5
            J += a * r / (1 + 0.8 * r)
    assert abs(O.Walker(sysd, pos).jastrow() - J) <= 1e-12 * abs(J)


def test_baseline_ewald_is_alpha_independent_within_its_truncation(oracle):
    """The Ewald energy does not depend on alpha once converged: with
    alpha = 8/L over every minimum-image pair and the reciprocal tolerance
    rule it is converged to 1e-6 (alpha = 7/L agrees), and the benchmark's
    truncated rules (alpha = 5/L, cutoff L/2, |n|^2 <= 25) stay below it by
    the mean-field size of what they leave out, over 20 uniform
    configurations of N = 54 at r_s = 2, within 4 standard errors.  A wrong
    self or background term or reciprocal weight misses by > 0.1 Ha."""
    O = oracle
    N, rs = 54, 2.0
    L = box_length(N, rs)
    base = unpolarised(O, N, rs, 0.0, 0.0, 1.0, baseline=True)
    conv = unpolarised(O, N, rs, 0.0, 0.0, 1.0, baseline=False)
    conv7 = O.System(27, 27, L, 0.0, 0.0, 1.0, 7.0 / L, 0, 0.0, 1)
    rng = np.random.default_rng(2024)
    diffs = []
    for _ in range(20):
        pos = rng.uniform(0, L, (N, 3))
        v_base = O.Walker(base, pos).local_energy()[2]
        v_conv = O.Walker(conv, pos).local_energy()[2]
        v_conv7 = O.Walker(conv7, pos).local_energy()[2]
        assert abs(v_conv - v_conv7) <= 1e-5, (v_conv, v_conv7)
        diffs.append(v_base - v_conv)
    diffs = np.array(diffs)
    expect = -ewald_truncation(N, L, 5.0 / L, 0.5 * L, 25)
    se = diffs.std(ddof=1) / np.sqrt(len(diffs))
    assert abs(diffs.mean() - expect) <= 4 * se + 1e-5, (diffs.mean(), expect, se)


def test_hartree_fock_closed_form():
    """E_HF per electron of N = 54 at r_s = 2: kinetic sum of k^2 / 2 (the
    zero-variance value pinned elsewhere), exchange and the Madelung term."""
    e, t, x, m = hartree_fock_energy(54, 2.0)
    assert abs(t / 54 - 0.2658550000) < 1e-9
    assert abs(e / 54 - 0.0187674) < 1e-7
    assert abs(x / 54 + 0.1306704) < 1e-7 and abs(m / 54 + 0.1164173) < 1e-7


def test_jastrow_free_chain_reproduces_the_hartree_fock_energy(oracle):
    """The oracle's unpolarised Gaussian sweep chain without a Jastrow factor
    (a_same = a_anti = 0) samples |D_up D_down|^2, whose mean local energy is
    the Hartree-Fock energy of the plane-wave determinant (less the
    benchmark Ewald rules' truncation): 16 chains of N = 54 at r_s = 2, 120
    measured sweeps each, within 4 blocking errors."""
    O = oracle
    N, rs = 54, 2.0
    sysd = unpolarised(O, N, rs, 0.0, 0.0, 1.0)
    L = sysd.L
    e_hf = hartree_fock_energy(N, rs)[0] - ewald_truncation(N, L, 5.0 / L, 0.5 * L, 25)
    chains, warm, meas, step = 16, 30, 120, 0.45
    means = []
    for c in range(chains):
        pos, _ = O.init_positions(sysd, 99, c)
        w = O.Walker(sysd, pos)
        s = step
        for sw in range(warm):
            w.sweep(99, c, sw, s, move_kind=1, adapt=False, measure=False)
        es = []
        for sw in range(warm, warm + meas):
            _, _, e = w.sweep(99, c, sw, s, move_kind=1, adapt=False, measure=True)
            es.append(e)
        means.append(np.mean(es))
    means = np.array(means) / N
    se = means.std(ddof=1) / np.sqrt(chains)
    assert abs(means.mean() - e_hf / N) <= 4 * se, (means.mean(), e_hf / N, se)
