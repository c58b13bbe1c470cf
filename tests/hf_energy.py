"""Closed forms for the Jastrow-free homogeneous electron gas in a periodic
box: the Hartree-Fock energy of the closed-shell plane-wave determinant and
the mean-field size of the truncation of the Ewald sum used by the
benchmark configurations (alpha = 5/L, real-space cutoff L/2, |n|^2 <= 25).
Independent of the oracle and of the GPU code (numpy and scipy only)."""
import numpy as np
from scipy.integrate import quad
from scipy.special import erfc

MADELUNG_SC = 2.837297479480619  # simple cubic, with neutralising background


def box_length(n_total, rs):
    return float(np.cbrt(4.0 * np.pi * n_total / 3.0) * rs)


def occupied_k(n_spin, L):
    """The n_spin lowest plane waves 2 pi n / L (closed shells: all +-n of
    each occupied |n|^2)."""
    R = 4
    ns = np.array([(x, y, z) for x in range(-R, R + 1) for y in range(-R, R + 1) for z in range(-R, R + 1)])
    n2 = (ns ** 2).sum(1)
    order = np.argsort(n2, kind="stable")
    sel = ns[order[:n_spin]]
    assert (sel ** 2).sum(1).max() < (ns[order[n_spin]] ** 2).sum(), "not a closed shell"
    return 2.0 * np.pi / L * sel


def hartree_fock_energy(n_total, rs):
    """(total, kinetic, exchange, Madelung) of the unpolarised closed-shell
    plane-wave determinant: T = sum 1/2 k^2, E_x = -(1/2V) sum_spin
    sum_{k != k'} 4 pi / |k - k'|^2, and each electron's interaction with its
    own images, N v_M / 2 with v_M = -2.837297 / L."""
    L = box_length(n_total, rs)
    V = L ** 3
    k = occupied_k(n_total // 2, L)
    T = 2 * 0.5 * float((k ** 2).sum())
    q2 = ((k[:, None, :] - k[None, :, :]) ** 2).sum(-1)
    off = q2 > 0
    Ex = -(1.0 / (2.0 * V)) * 2 * float((4.0 * np.pi / q2[off]).sum())
    mad = n_total * (-MADELUNG_SC / L) / 2.0
    return T + Ex + mad, T, Ex, mad


def ewald_truncation(n_total, L, alpha, rcut, nmax2):
    """Mean-field size of what the truncated Ewald sum leaves out for a
    uniform pair density: the real-space pairs beyond rcut,
    (N^2 / 2V) 4 pi int_rcut^inf r erfc(alpha r) dr, and the reciprocal
    terms beyond |n|^2 = nmax2 with <|S_G|^2> = N,
    (N / 2 pi V) sum_{half space} 8 pi^2 / G^2 exp(-G^2 / 4 alpha^2).
    The truncated sum is lower than the converged one by about their sum."""
    V = L ** 3
    I = quad(lambda x: x * erfc(x), alpha * rcut, np.inf)[0] / alpha ** 2
    real = n_total ** 2 / (2 * V) * 4 * np.pi * I
    R = 12
    ns = np.array([(x, y, z) for x in range(-R, R + 1) for y in range(-R, R + 1) for z in range(-R, R + 1)])
    n2 = (ns ** 2).sum(1)
    half = (ns[:, 0] > 0) | ((ns[:, 0] == 0) & (ns[:, 1] > 0)) | ((ns[:, 0] == 0) & (ns[:, 1] == 0) & (ns[:, 2] > 0))
    sel = half & (n2 > nmax2)
    G2 = (2 * np.pi / L) ** 2 * n2[sel]
    recip = n_total / (2 * np.pi * V) * float((8 * np.pi ** 2 / G2 * np.exp(-G2 / (4 * alpha ** 2))).sum())
    return real + recip
