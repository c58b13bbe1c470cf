"""Physics and statistics of the device chains against closed forms and the
oracle: the per-walker blocking state that feeds every reported standard
error, the Jastrow-free chain's Hartree-Fock energy on the timed kernels,
and the mixed and FP64 paths' energies from independent seeds."""
import ctypes as C

import numpy as np
import pytest

from helpers import oracle_system
from hf_energy import ewald_truncation, hartree_fock_energy

pytestmark = pytest.mark.gpu


def oracle_blocking_state(O, block_size, xs):
    b = O.Blocking()
    O.lib().orc_blocking_init(C.byref(b), block_size)
    for x in xs:
        O.lib().orc_blocking_add(C.byref(b), float(x))
    return b


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_device_blocking_matches_the_oracle_blocking(vmc, oracle, precision):
    """VMC_STAT_BLK_* after 23 measured sweeps in blocks of 3 equal the
    oracle's blocking analysis (blocking_analysis.cpp:3-41 restated) of each
    walker's local-energy stream as the device measured it, and the walker
    aggregate's standard error is sqrt(sum SE^2) / W of those
    (main.cpp:79-104).  In FP64 the stream is also the oracle twin's."""
    s = vmc.System.heg(54, 2.0, 1.0, precision=precision, reinvert_interval=64 if precision else 0)
    W, seed, bs, sweeps, step = 4, 31337, 3, 23, 0.6
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    twins = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    streams = [[] for _ in range(W)]
    for sw in range(sweeps):
        ens.sweep(1, measure=True, block_size=bs)
        st = ens.stats()
        for w in range(W):
            streams[w].append(st[w, vmc.STAT["LAST_E"]])
            if precision == 0:
                _, _, e = twins[w].sweep(seed, w, sw, step, move_kind=s.move_kind, adapt=False, measure=True)
                assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N)
    st = ens.stats()
    se2 = 0.0
    for w in range(W):
        b = oracle_blocking_state(oracle, bs, streams[w])
        assert st[w, vmc.STAT["BLK_N"]] == b.num_blocks == sweeps // bs
        assert st[w, vmc.STAT["BLK_IN"]] == b.in_block == sweeps % bs
        assert abs(st[w, vmc.STAT["BLK_MEAN"]] - b.running_mean) <= 1e-12 * abs(b.running_mean)
        assert abs(st[w, vmc.STAT["BLK_M2"]] - b.running_m2) <= 1e-9 * abs(b.running_m2)
        assert abs(st[w, vmc.STAT["BLK_SUM"]] - b.block_sum) <= 1e-12 * max(abs(b.block_sum), 1.0)
        m, se = oracle.blocking(bs, streams[w])
        se2 += se * se
    mean, se = vmc.blocking_summary(st)
    assert abs(se - np.sqrt(se2) / W) <= 1e-12 * se


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_jastrow_free_chain_reproduces_hartree_fock(vmc, precision):
    """Without a Jastrow factor the chain samples |D_up D_down|^2, whose mean
    local energy is the Hartree-Fock energy of the plane-wave determinant,
    0.0187674 Ha per electron at N = 54, r_s = 2, less the benchmark Ewald
    rules' truncation (mean-field estimate, 2.2e-4 Ha per electron): the
    register kernels with 4096 walkers, 300 measured sweeps, within 4
    blocking errors plus a tenth of the truncation."""
    N, rs = 54, 2.0
    s = vmc.System.heg(N, rs, 1.0, precision=precision, reinvert_interval=256 if precision else 0,
                       a_same=0.0, a_anti=0.0)
    L = s.box_length
    cut = ewald_truncation(N, L, 5.0 / L, 0.5 * L, 25)
    expect = (hartree_fock_energy(N, rs)[0] - cut) / N
    ens = vmc.Ensemble(s, 4096, 77 + precision, step0=0.45)
    assert ens.kernel() == vmc.KERNEL_REGISTER
    ens.init_positions()
    ens.sweep(60, adapt=True)
    ens.reset_stats()
    ens.sweep(300, measure=True, block_size=30)
    mean, se = vmc.blocking_summary(ens.stats())
    e = mean / N
    print(f"E/N {e:.7f} +- {se / N:.2e}, Hartree-Fock (truncated Ewald) {expect:.7f}", flush=True)
    assert abs(e - expect) <= 4 * se / N + 0.1 * cut / N


def test_mixed_and_fp64_energies_agree_from_independent_seeds(vmc):
    """N = 54, r_s = 2, b = 1 with 4096 walkers: the mixed-precision and the
    FP64 paths from different seeds give the same energy per electron
    within 3 combined blocking errors."""
    out = {}
    for prec, seed in ((0, 11), (1, 12)):
        s = vmc.System.heg(54, 2.0, 1.0, precision=prec, reinvert_interval=256 if prec else 0)
        ens = vmc.Ensemble(s, 4096, seed, step0=0.6)
        ens.init_positions()
        ens.sweep(100, adapt=True)
        ens.reset_stats()
        ens.sweep(500, measure=True, block_size=50)
        out[prec] = vmc.blocking_summary(ens.stats())
    (m0, s0), (m1, s1) = out[0], out[1]
    print(f"FP64 {m0 / 54:.7f}({s0 / 54:.1e}) mixed {m1 / 54:.7f}({s1 / 54:.1e})", flush=True)
    assert abs(m0 - m1) <= 3 * np.hypot(s0, s1)
