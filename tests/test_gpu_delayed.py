"""GPU parity of the FP64 delayed-update sweep (vmc_sweep_du.hip).

The delayed-update kernel keeps each determinant's FP64 inverse in HBM and
applies accepted moves in windows of eight as Woodbury updates on the matrix
cores; the reference updates the whole inverse on every accepted move
(slater_plane_wave.cpp:367-423).  In FP64 both are the same chain: the
device follows the oracle's Philox twin move for move, with the local energy
after every sweep within 1e-9 of the oracle's at the same positions.
"""
import numpy as np
import pytest

from helpers import oracle_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,rs,W,sweeps,force", [(54, 2.0, 4, 6, True), (38, 1.0, 4, 6, True),
                                                  (246, 2.0, 2, 2, False)])
def test_delayed_chain_follows_the_oracle_twin(vmc, oracle, N, rs, W, sweeps, force):
    s = vmc.System.heg(N, rs, 1.0)
    seed, step = 500 + N, 0.4 * s.box_length / N ** (1 / 3)
    ens = vmc.Ensemble(s, W, seed, step0=step)
    if force:
        ens.set_kernel(vmc.KERNEL_DELAYED)
    assert ens.kernel() == vmc.KERNEL_DELAYED
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    for sweep in range(sweeps):
        ens.sweep(1, measure=True, adapt=False)
        st = ens.stats()
        gpu_pos = ens.positions()
        print(f"N={N} sweep {sweep}", flush=True)
        for w in range(W):
            acc, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                         measure=True)
            np.testing.assert_allclose(gpu_pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N), (w, sweep)
    assert np.all(st[:, vmc.STAT["ACCEPTED"]] > 0)


def test_delayed_inverse_after_many_sweeps(vmc, oracle):
    """N = 294 (147 + 147) in FP64 over 30 adaptive sweeps of 64 walkers:
    the inverse the windows leave in HBM equals a fresh FP64 inversion at the
    final positions to 1e-9 of its largest element, and log|det| tracked
    through the accepted ratios matches the fresh one."""
    s = vmc.System.heg(294, 2.0, 1.0)
    W = 64
    ens = vmc.Ensemble(s, W, 2940, step0=0.3)
    assert ens.kernel() == vmc.KERNEL_DELAYED
    ens.init_positions()
    ens.sweep(30, adapt=True)
    inv = ens.inverse()
    pos = ens.positions()
    parts = ens.measure()
    ens.close()
    osys = oracle_system(oracle, s)
    n = s.n_up
    worst = 0.0
    for w in range(0, W, 8):
        wk = oracle.Walker(osys, pos[w])
        for d in range(2):
            ref = wk.inverse(d)
            got = inv[w, d * n * n:(d + 1) * n * n].reshape(n, n)
            worst = max(worst, float(np.abs(got - ref).max() / np.abs(ref).max()))
            assert abs(parts[f"LOGDET{d}"][w] - wk.logdet(d)) <= 1e-9 * max(abs(wk.logdet(d)), 1.0)
        e, _, _ = wk.local_energy()
        assert abs(parts["ELOCAL"][w] - e) <= 1e-10 * max(abs(e), 1e-3 * s.N)
    print(f"inverse after 30 sweeps: worst {worst:.1e}", flush=True)
    assert worst <= 1e-9
