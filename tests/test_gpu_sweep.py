"""GPU sweep against its CPU twin.

The oracle's ensemble driver (orc_sweep) draws the same Philox streams as the
device sweep, so an FP64 GPU chain and the oracle chain see identical
proposals and accept uniforms.  Their trajectories therefore agree to
rounding until an accept test lands within rounding of its threshold, which
over a few sweeps does not happen for these seeds: positions, accept counts
and local energies are compared move for move.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# The mixed-precision local energy against the FP64 oracle at the same
# configuration: within 1e-6 of max(|E_L|, N x 1 mHa) (it is evaluated in FP64
# from an inverse refined to FP64 accuracy).
MIXED_E_L = 1e-6


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


KERNELS = [("generic", 1, 0), ("register", 2, 0), ("register_o4", 2, 1)]


@pytest.mark.parametrize("kname,kernel,variant", KERNELS, ids=[k[0] for k in KERNELS])
@pytest.mark.parametrize("name,factory", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_fp64_sweeps_follow_oracle_twin(vmc, oracle, name, factory, kname, kernel, variant):
    s = factory(vmc)
    W, seed, sweeps = 4, 987654321, 6
    step = 0.3 * s.box_length / 10.0 * 3.0
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.set_kernel(kernel)
    ens.set_rr_variant(variant)
    assert ens.kernel() == kernel
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


@pytest.mark.parametrize("kernel", [1, 2], ids=["generic", "register"])
def test_adaptation_matches_reference_rule(vmc, oracle, kernel):
    """Warmup step rule s *= exp(0.25 (acc - 0.5)), s <= L/2 (simulation.cpp:178-194)."""
    s = vmc.System.heg(14, 1.0, 1.0)
    W, seed = 3, 5
    ens = vmc.Ensemble(s, W, seed, step0=0.05)
    ens.set_kernel(kernel)
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


@pytest.mark.parametrize("kernel", [1, 2], ids=["generic", "register"])
def test_zero_step_accepts_everything_and_keeps_energy(vmc, kernel):
    """Zero-step proposals are always accepted and leave E_L unchanged
    (test_simulation.cpp:91-109)."""
    s = vmc.System.reference(7, 6.0)
    ens = vmc.Ensemble(s, 8, 314159, step0=0.0)
    ens.set_kernel(kernel)
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


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_register_and_generic_kernels_agree(vmc, prec):
    """Both sweep kernels on the same Philox streams: FP64 chains coincide to
    rounding; mixed chains (different FP32 operation orders) stay statistically
    equal and their local energies agree to the mixed budget while the
    trajectories coincide."""
    s = vmc.System.heg(54, 2.0, 1.0, precision=prec, reinvert_interval=64 if prec else 0)
    W, seed = 64, 4242
    out = []
    for kernel in (1, 2):
        ens = vmc.Ensemble(s, W, seed, step0=0.6)
        ens.set_kernel(kernel)
        ens.init_positions()
        ens.sweep(3, measure=True)
        out.append((ens.positions(), ens.stats()))
    (p1, s1), (p2, s2) = out
    if prec == 0:
        np.testing.assert_allclose(p1, p2, atol=1e-9)
        np.testing.assert_allclose(s1[:, vmc.STAT["LAST_E"]], s2[:, vmc.STAT["LAST_E"]], rtol=1e-9,
                                   atol=1e-9)
        assert np.array_equal(s1[:, vmc.STAT["ACCEPTED"]], s2[:, vmc.STAT["ACCEPTED"]])
    else:
        close = np.abs(p1 - p2).reshape(W, -1).max(1) < 1e-4
        assert close.mean() > 0.8
        e1, e2 = s1[close, vmc.STAT["LAST_E"]], s2[close, vmc.STAT["LAST_E"]]
        assert np.max(np.abs(e1 - e2) / np.maximum(np.abs(e1), 1.0)) < 1e-4


def test_fp64_reinversion_keeps_the_oracle_twin(vmc, oracle):
    """The post-sweep FP64 re-inversion (Newton-Schulz on the matrix cores)
    replaces an inverse that the reference keeps by Sherman-Morrison only
    (slater_plane_wave.cpp accept path); in FP64 both are the same matrix to
    rounding, so a chain that re-inverts every 16 accepted moves still follows
    the oracle twin move for move, and the drift found before each
    re-inversion is FP64 rounding."""
    s = vmc.System.heg(54, 2.0, 1.0, reinvert_interval=16)
    W, seed, sweeps = 4, 77, 6
    step = 0.6
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.set_kernel(2)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    for sweep in range(sweeps):
        ens.sweep(1, measure=True, adapt=False)
        st = ens.stats()
        gpu_pos = ens.positions()
        for w in range(W):
            _, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                       measure=True)
            np.testing.assert_allclose(gpu_pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= 1)
    d = ens.drift()
    assert np.all(d > 0) and np.all(d < 1e-11), d


def test_mixed_reinversion_bounds_the_drift(vmc, oracle):
    """Mixed precision: FP32 Sherman-Morrison drift is measured at every FP64
    re-inversion (max |I - D^-1 D| over the K accepted moves before it) and
    stays far below the mixed-precision budget; the local energy after many
    sweeps still matches the FP64 oracle at the walker's own positions to
    that budget (relative 1e-5 of the sum of the magnitudes of its parts)."""
    s = vmc.System.heg(54, 2.0, 1.0, precision=1, reinvert_interval=64)
    W, seed = 64, 2026
    ens = vmc.Ensemble(s, W, seed, step0=0.6)
    ens.set_kernel(2)
    ens.init_positions()
    ens.sweep(40, measure=True, adapt=False)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= 5)
    d = ens.drift()
    assert np.all(d > 0) and np.all(d < 1e-3), d.max()
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(0, W, 8):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= MIXED_E_L * max(abs(e), 1e-3 * s.N), (w, st[w, vmc.STAT["LAST_E"]], e)


def test_kernel_timing_counts_every_launch(vmc):
    """The per-kernel event timer sees each sweep, re-inversion and
    local-energy launch of the register path, and forgets them once read."""
    s = vmc.System.heg(54, 2.0, 1.0, precision=1, reinvert_interval=64)
    ens = vmc.Ensemble(s, 32, 9, step0=0.6)
    ens.set_kernel(2)
    ens.init_positions()
    ens.set_timing(True)
    ens.sweep(3, measure=True)
    ens.sweep(2, measure=False)
    t = ens.timing()
    # (a mixed measured sweep's local energy refreshes every inverse itself:
    # only the two unmeasured sweeps launch the re-inversion)
    assert t["sweep"][1] == 5 and t["reinvert"][1] == 2 and t["measure"][1] == 3
    assert all(v[0] > 0 for v in t.values())
    assert all(v[1] == 0 for v in ens.timing().values())


@pytest.mark.parametrize("K", [16, 64, 256])
def test_n114_reinversion_interval_drift_and_energy(vmc, oracle, K):
    """N = 114, r_s = 5, b = 0.8 in mixed precision with FP64 re-inversion
    every K accepted moves (the generic kernel): the inverse drift met before
    each re-inversion is recorded, stays small, and the local energy after
    the sweeps matches the FP64 oracle at the walker's positions."""
    s = vmc.System.heg(114, 5.0, 0.8, precision=1, reinvert_interval=K)
    W = 16
    ens = vmc.Ensemble(s, W, 1140 + K, step0=1.5)
    ens.init_positions()
    ens.sweep(12, measure=True, adapt=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= (1 if K == 256 else 4))
    d = ens.drift()
    assert np.all(d > 0) and np.all(d < 1e-3), d
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(0, W, 4):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= MIXED_E_L * max(abs(e), 1e-3 * s.N), (w, st[w, vmc.STAT["LAST_E"]], e)


@pytest.mark.parametrize("threads", [128, 256])
def test_generic_multi_wavefront_walkers_follow_the_oracle_twin(vmc, oracle, threads):
    """The generic kernel with 2 or 4 wavefronts per walker (workgroup-wide
    sums) follows the oracle's FP64 twin move for move, as the one-wavefront
    kernel does; N = 66 (33 + 33, no register instantiation)."""
    s = vmc.System.heg(66, 2.0, 1.0)
    W, seed, sweeps, step = 3, 6606, 3, 0.6
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.set_walker_threads(threads)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    for sweep in range(sweeps):
        ens.sweep(1, measure=True)
        st, pos = ens.stats(), ens.positions()
        for w in range(W):
            _, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                       measure=True)
            np.testing.assert_allclose(pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N)


def test_generic_multi_wavefront_mixed_with_reinversion(vmc, oracle):
    """N = 162 (81 + 81, inverse in HBM), 4 wavefronts per walker, mixed
    precision with FP64 re-inversion every 64 accepted moves: drift recorded,
    local energy within the mixed budget of the FP64 oracle."""
    s = vmc.System.heg(162, 3.0, 1.0, precision=1, reinvert_interval=64)
    ens = vmc.Ensemble(s, 4, 162, step0=0.9)
    ens.set_walker_threads(256)
    ens.init_positions()
    ens.sweep(4, measure=True, adapt=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= 2)
    assert np.all(ens.drift() > 0) and np.all(ens.drift() < 1e-3)
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(4):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= MIXED_E_L * max(abs(e), 1e-3 * s.N)


RRW_SYSTEMS = [
    ("heg_N66_rs2", lambda V: V.System.heg(66, 2.0, 1.0)),
    ("heg_N114_rs5", lambda V: V.System.heg(114, 5.0, 0.8)),
    ("heg_N162_rs3", lambda V: V.System.heg(162, 3.0, 1.2)),
    ("ref_N57", lambda V: V.System.reference(57, 12.0)),
]


@pytest.mark.parametrize("name,factory", RRW_SYSTEMS, ids=[s[0] for s in RRW_SYSTEMS])
def test_multiwave_register_kernel_follows_oracle_twin(vmc, oracle, name, factory):
    """Register-resident rows with one wavefront per spin (two for 81
    electrons per spin) follow the oracle's FP64 twin move for move, like the
    other kernels."""
    s = factory(vmc)
    W, seed, sweeps = 2, 31337, 2
    step = 0.05 * s.box_length
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.set_kernel(vmc.KERNEL_REGISTER)
    assert ens.kernel() == vmc.KERNEL_REGISTER
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    acc = np.zeros(W)
    for sweep in range(sweeps):
        ens.sweep(1, measure=True)
        st, pos = ens.stats(), ens.positions()
        for w in range(W):
            a, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                       measure=True)
            acc[w] += a
            np.testing.assert_allclose(pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert st[w, vmc.STAT["ACCEPTED"]] == acc[w]
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N)


def test_multiwave_register_kernel_mixed_reinversion(vmc, oracle):
    """N = 114 mixed precision on the multi-wavefront register kernel with
    FP64 re-inversion every 64 accepted moves: drift recorded, local energy
    within the mixed budget of the FP64 oracle, and statistically the same
    chain as the generic kernel."""
    s = vmc.System.heg(114, 5.0, 0.8, precision=1, reinvert_interval=64)
    W = 32
    ens = vmc.Ensemble(s, W, 57, step0=1.5)
    ens.set_kernel(vmc.KERNEL_REGISTER)
    ens.init_positions()
    ens.sweep(8, measure=True, adapt=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= 3)
    d = ens.drift()
    assert np.all(d > 0) and np.all(d < 1e-3), d.max()
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(0, W, 8):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= MIXED_E_L * max(abs(e), 1e-3 * s.N)


@pytest.mark.parametrize("N", [246, 294])
def test_multiwave_register_kernel_large_spins_mixed(vmc, oracle, N):
    """123 and 147 electrons per spin (two and three wavefronts per spin, FP32
    rows in registers): after sweeps with FP64 re-inversion the local energy
    matches the FP64 oracle at the walker's positions to the mixed budget."""
    s = vmc.System.heg(N, 4.0, 1.0, precision=1, reinvert_interval=128)
    W = 4
    ens = vmc.Ensemble(s, W, N, step0=1.2)
    ens.set_kernel(vmc.KERNEL_REGISTER)
    assert ens.kernel() == vmc.KERNEL_REGISTER
    ens.init_positions()
    ens.sweep(3, measure=True, adapt=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] >= 1)
    assert np.all(ens.drift() < 1e-3)
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    for w in range(W):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= MIXED_E_L * max(abs(e), 1e-3 * s.N), (w, st[w, vmc.STAT["LAST_E"]], e)


NS_SYSTEMS = [
    ("heg_N66_rs2_register", lambda V: V.System.heg(66, 2.0, 1.0, precision=1, reinvert_interval=16), 2),
    ("heg_N114_rs5_register", lambda V: V.System.heg(114, 5.0, 0.8, precision=1, reinvert_interval=16), 2),
    ("heg_N162_rs3_register", lambda V: V.System.heg(162, 3.0, 1.2, precision=1, reinvert_interval=16), 2),
    ("heg_N246_rs4_register", lambda V: V.System.heg(246, 4.0, 1.0, precision=1, reinvert_interval=16), 2),
    ("heg_N294_rs4_register", lambda V: V.System.heg(294, 4.0, 1.0, precision=1, reinvert_interval=16), 2),
    ("heg_N114_rs5_generic", lambda V: V.System.heg(114, 5.0, 0.8, precision=1, reinvert_interval=16), 1),
    ("heg_N246_rs4_generic", lambda V: V.System.heg(246, 4.0, 1.0, precision=1, reinvert_interval=16), 1),
    ("ref_N57_generic", lambda V: dataclasses.replace(V.System.reference(57, 12.0, precision=1),
                                                      reinvert_interval=16), 1),
]


@pytest.mark.parametrize("name,factory,kernel", NS_SYSTEMS, ids=[s[0] for s in NS_SYSTEMS])
def test_matrix_core_reinversion_reaches_fp32_rounding(vmc, oracle, name, factory, kernel):
    """FP32 inverses of 33 to 147 electrons per spin are re-inverted by
    Newton-Schulz steps on the FP64 matrix cores (operands in LDS up to 80 per
    spin, E and the residual through the caches above).  Every measured sweep
    re-inverts every walker, its last step's FP64 result feeding the FP64
    local energy, so the local energy's error against the FP64 oracle shows
    the quality of the re-inverted inverse: within 1e-6 of max(|E_L|,
    N x 1 mHa).  A wrong or unconverged inverse (drift 1e-4 or more) misses
    this bound."""
    s = factory(vmc)
    W = 8
    ens = vmc.Ensemble(s, W, 4242, step0=0.3 * s.box_length / 6)
    ens.set_kernel(kernel)
    assert ens.kernel() == kernel
    ens.init_positions()
    ens.sweep(3, measure=True, adapt=True)
    st = ens.stats()
    assert np.all(st[:, vmc.STAT["REINV"]] == 3), st[:, vmc.STAT["REINV"]]
    d = ens.drift()
    assert np.all(d > 0) and np.all(d < 1e-3), d
    pos = ens.positions()
    osys = oracle_system(oracle, s)
    worst = 0.0
    for w in range(W):
        wk = oracle.Walker(osys, pos[w])
        e, kin, _ = wk.local_energy()
        worst = max(worst, abs(st[w, vmc.STAT["LAST_E"]] - e) / max(abs(e), 1e-3 * s.N))
    print(f"{name}: E_L relative error max {worst:.2e}, drift max {d.max():.2e}")
    assert worst <= MIXED_E_L, worst


@pytest.mark.parametrize("name,factory", [("heg_N54_rs2", lambda V: V.System.heg(54, 2.0, 1.0)),
                                          ("heg_N114_rs5", lambda V: V.System.heg(114, 5.0, 0.8))],
                         ids=["heg_N54_rs2", "heg_N114_rs5"])
def test_kernel_switches_keep_the_oracle_twin(vmc, oracle, name, factory):
    """The register kernels keep each walker's Jastrow row sums between
    sweeps (recomputed every 16 sweeps); a sweep of the generic kernel in
    between moves the electrons without them and marks them stale.  A chain
    that switches kernels (register, generic, register ...) over 20 sweeps,
    across the 16-sweep recomputation, follows the oracle's FP64 twin move
    for move."""
    s = factory(vmc)
    W, seed = 3, 24680
    step = 0.3 * s.box_length / 10.0 * 3.0
    ens = vmc.Ensemble(s, W, seed, step0=step)
    ens.init_positions()
    osys = oracle_system(oracle, s)
    walkers = [oracle.Walker(osys, oracle.init_positions(osys, seed, w)[0]) for w in range(W)]
    plan = [2, 2, 1, 2, 2, 2, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 2, 2]
    for sweep, kernel in enumerate(plan):
        ens.set_kernel(kernel)
        ens.sweep(1, measure=True, adapt=False)
        st, pos = ens.stats(), ens.positions()
        for w in range(W):
            _, _, e = walkers[w].sweep(seed, w, sweep, step, move_kind=s.move_kind, adapt=False,
                                       measure=True)
            np.testing.assert_allclose(pos[w], walkers[w].positions(), rtol=0, atol=1e-9)
            assert abs(st[w, vmc.STAT["LAST_E"]] - e) <= 1e-9 * max(abs(e), 1e-3 * s.N), (sweep, w)


@pytest.mark.parametrize("prec,W,groups", [(1, 2048, (1, 4, 3)), (0, 1030, (1, 2, 5)), (1, 300, (1, 8))])
def test_walker_groups_give_the_same_chains(vmc, prec, W, groups):
    """Measured register sweeps split over walker groups on their own streams
    (ragged splits included) leave every walker exactly where one group
    leaves it: positions, inverses and statistics bit for bit."""
    out = []
    for g in groups:
        s = vmc.System.heg(54, 2.0, 1.0, precision=prec, reinvert_interval=64 if prec else 0)
        e = vmc.Ensemble(s, W, 77, step0=0.6)
        e.set_walker_groups(g)
        e.init_positions()
        e.sweep(2, adapt=True)
        e.sweep(4, measure=True, block_size=2)
        e.sweep(1, measure=True, block_size=2)
        out.append((e.positions(), e.inverse(), e.stats()))
        print(f"prec={prec} W={W} groups={g} done", flush=True)
        e.close()
    for g, (p, i, st) in zip(groups[1:], out[1:]):
        np.testing.assert_array_equal(p, out[0][0], err_msg=f"groups={g}")
        np.testing.assert_array_equal(i, out[0][1], err_msg=f"groups={g}")
        np.testing.assert_array_equal(st, out[0][2], err_msg=f"groups={g}")
