"""GPU sweep against its CPU twin.

The oracle's ensemble driver (orc_sweep) draws the same Philox streams as the
device sweep, so an FP64 GPU chain and the oracle chain see identical
proposals and accept uniforms.  Their trajectories therefore agree to
rounding until an accept test lands within rounding of its threshold, which
over a few sweeps does not happen for these seeds: positions, accept counts
and local energies are compared move for move.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_system(O, s):
    return O.System(s.n_up, s.n_dn, s.box_length, s.a_same, s.a_anti, s.b, s.alpha,
                    s.recip_nmax2, s.real_cutoff, s.erfc_kind)


SYSTEMS = [
    ("heg_N14_rs1", lambda V: V.System.heg(14, 1.0, 1.0)),
    ("heg_N54_rs2", lambda V: V.System.heg(54, 2.0, 1.0)),
    ("ref_N19_uniform", lambda V: V.System.reference(19, 8.0)),
]


@pytest.mark.parametrize("name,factory", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_init_positions_match_oracle(vmc, oracle, name, factory):
    s = factory(vmc)
    W, seed = 6, 0x1234ABCD5678
    ens = vmc.Ensemble(s, W, seed)
    ens.init_positions()
    gpu = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(W):
        pos, att = oracle.init_positions(osys, seed, w)
        assert att >= 0
        np.testing.assert_array_equal(gpu[w], pos)


@pytest.mark.parametrize("name,factory", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_fp64_sweeps_follow_oracle_twin(vmc, oracle, name, factory):
    s = factory(vmc)
    W, seed, sweeps = 4, 987654321, 6
    step = 0.3 * s.box_length / 10.0 * 3.0
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = []
    for w in range(W):
        pos, _ = oracle.init_positions(osys, seed, w)
        walkers.append(oracle.Walker(osys, pos))
    acc_total = np.zeros(W)
    for sweep in range(sweeps):
        ens.sweep(1, measure=True, adapt=False)
        st = ens.stats()
        gpu_pos = ens.positions()
        for w in range(W):
            acc, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                         measure=True)
            acc_total[w] += acc
            np.testing.assert_allclose(gpu_pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert st[w, vmc.STAT["ACCEPTED"]] == acc_total[w]
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N)
    assert np.all(ens.stats()[:, vmc.STAT["PROPOSED"]] == sweeps * s.N)


def test_adaptation_matches_reference_rule(vmc, oracle):
    """Warmup step rule s *= exp(0.25 (acc - 0.5)), s <= L/2 (simulation.cpp:178-194)."""
    s = vmc.System.heg(14, 1.0, 1.0)
    W, seed = 3, 5
    ens = vmc.Ensemble(s, W, seed, step0=0.05)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    steps = [0.05] * W
    for sweep in range(4):
        ens.sweep(1, measure=False, adapt=True)
        for w in range(W):
            _, steps[w], _ = walkers[w].sweep(seed, w, sweep, steps[w], move_kind=s.move_kind,
                                              adapt=True, measure=False)
    np.testing.assert_allclose(ens.steps(), steps, rtol=1e-12)


def test_zero_step_accepts_everything_and_keeps_energy(vmc):
    """Zero-step proposals are always accepted and leave E_L unchanged
    (test_simulation.cpp:91-109)."""
    s = vmc.System.reference(7, 6.0)
    ens = vmc.Ensemble(s, 8, 314159, step0=0.0)
    ens.init_positions()
    ens.sweep(1, measure=True)
    e0 = ens.stats()[:, vmc.STAT["LAST_E"]].copy()
    ens.sweep(3, measure=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["ACCEPTED"]] == 4 * 7)
    np.testing.assert_allclose(st[:, vmc.STAT["LAST_E"]], e0, rtol=1e-11, atol=1e-11)


def test_large_step_rejects(vmc):
    """A large proposal step produces rejections (test_simulation.cpp:111-122)."""
    s = vmc.System.reference(7, 6.0)
    ens = vmc.Ensemble(s, 8, 20250308, step0=2.0)
    ens.init_positions()
    ens.sweep(15)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["ACCEPTED"]] < st[:, vmc.STAT["PROPOSED"]])
