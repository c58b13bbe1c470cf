"""Uneven batches: many systems' walker ensembles swept together, optionally
captured into a HIP graph.

Register-resident systems of one size form a group swept by one launch per
kernel; the systems on the generic kernel share one persistent launch.  A
walker's chain depends only on its system, seed, walker index, sweep index
and kernel, so every system of a batch must come out exactly as when it runs
alone with the same kernel -- positions, accept counts, step sizes and local
energies -- whatever the grouping, the stream overlap or graph replay.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # (N, r_s, b, walkers, seed)
    (14, 1.0, 1.0, 17, 101),
    (38, 3.0, 1.5, 9, 202),
    (66, 2.0, 0.8, 5, 303),
    (54, 5.0, 0.5, 12, 404),
    (186, 4.0, 1.0, 3, 505),   # 93 per spin: generic kernel only
    (54, 2.0, 1.2, 7, 606),
]


def systems(V, precision=0, K=0):
    return [V.System.heg(N, rs, b, precision=precision, reinvert_interval=K) for N, rs, b, _, _ in SPECS]


def snapshot(b):
    out = []
    for i in range(len(SPECS)):
        m = b.member(i)
        out.append((m.positions(), m.stats(), m.steps()))
    return out


def assert_same(a, b):
    for (p1, s1, t1), (p2, s2, t2) in zip(a, b):
        np.testing.assert_array_equal(p1, p2)
        np.testing.assert_array_equal(s1, s2)
        np.testing.assert_array_equal(t1, t2)


def alone(V, s, W, seed, kernel):
    """The system alone with the kernel its batch group runs: a batch takes
    the register kernels or the generic one, so an FP64 system that alone
    would take the delayed-update kernel runs the generic kernel there."""
    e = V.Ensemble(s, W, seed)
    e.set_kernel(kernel)
    if e.kernel() == V.KERNEL_DELAYED:
        e.set_kernel(V.KERNEL_GENERIC)
    e.init_positions()
    e.sweep(2, adapt=True, block_size=2)
    e.sweep(3, measure=True, block_size=2)
    out = (e.positions(), e.stats(), e.steps())
    e.close()
    return out


@pytest.mark.parametrize("precision,K", [(0, 0), (1, 32)], ids=["fp64", "mixed_K32"])
@pytest.mark.parametrize("kernel", [0, 1], ids=["auto", "generic"])
def test_batch_equals_each_system_alone(vmc, precision, K, kernel):
    sy = systems(vmc, precision, K)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    b = vmc.Batch(sy, W, seeds)
    b.set_kernel(kernel)
    info = b.info()
    fam, grp = b.groups()
    print("batch", info, fam, grp, flush=True)
    assert info["total_walkers"] == sum(W) and 1 <= info["grid"] <= sum(W)
    if kernel == 0:
        # 7, 19, 27 per spin: register; 33: multi-wavefront; 93: generic.
        # The two N = 54 systems share a group.
        assert sorted(fam) == [0, 1, 1, 1, 2]
        assert grp[3] == grp[5] and len(set(grp)) == 5
        assert fam[grp[4]] == 0 and fam[grp[2]] == 2 and fam[grp[0]] == 1
    else:
        assert fam == [0] and set(grp) == {0}
    b.init_positions()
    b.sweep(2, adapt=True, block_size=2)
    b.sweep(3, measure=True, block_size=2)
    got = snapshot(b)
    b.close()
    for i, (spec, s) in enumerate(zip(SPECS, sy)):
        ref = alone(vmc, s, spec[3], spec[4], kernel)
        assert_same([got[i]], [ref])
        assert np.all(ref[1][:, vmc.STAT["NMEAS"]] == 3)


def test_batch_overlap_and_placement_do_not_change_results(vmc):
    """Groups on their own streams or all on one, and the generic group's
    inverses in LDS or HBM: identical chains."""
    sy = systems(vmc, precision=1, K=64)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    outs = []
    for overlap, placement in ((True, None), (False, None), (True, [2] * len(SPECS))):
        b = vmc.Batch(sy, W, seeds, placement=placement)
        b.set_overlap(overlap)
        b.init_positions()
        b.sweep(3, measure=True, adapt=True)
        outs.append(snapshot(b))
        b.close()
    assert_same(outs[0], outs[1])
    assert_same(outs[0], outs[2])


@pytest.mark.parametrize("overlap", [True, False], ids=["overlap", "one_stream"])
def test_batch_graph_replays_continue_the_chains(vmc, overlap):
    """A captured graph of 2 sweeps (groups as parallel branches with
    overlap), replayed 3 times with eager reads between the replays, equals
    6 eager batch sweeps after every replay."""
    import torch

    sy = systems(vmc, precision=1, K=64)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    stream = torch.cuda.Stream()
    eager = vmc.Batch(sy, W, seeds, stream=stream.cuda_stream)
    graph = vmc.Batch(sy, W, seeds, stream=stream.cuda_stream)
    graph.set_overlap(overlap)
    eager.init_positions()
    graph.init_positions()
    graph.capture(2, measure=True, block_size=2)
    for rep in range(3):
        eager.sweep(2, measure=True, block_size=2)
        graph.replay()
        torch.cuda.synchronize()
        assert_same(snapshot(graph), snapshot(eager))
    for i in range(len(SPECS)):
        assert graph.member(i).sweeps_done() == 6
    eager.close()
    graph.close()


def test_batch_regroup_drops_the_graph(vmc):
    sy = systems(vmc, precision=1, K=64)[:2]
    b = vmc.Batch(sy, [4, 4], [1, 2])
    b.init_positions()
    b.capture(1)
    b.replay()
    b.set_kernel(vmc.KERNEL_GENERIC, system=1)
    fam, grp = b.groups()
    assert sorted(fam) == [0, 1] and fam[grp[1]] == 0
    with pytest.raises(vmc.VmcError):
        b.replay()
    b.close()


def test_reinversion_list_with_more_items_than_workgroups(vmc):
    """4000 walkers of N=14 at K=8 per determinant: a sweep accepts about 7
    moves per walker (3.5 per determinant), so well over a third of the
    walkers are due after each sweep and the lone
    ensemble's re-inversion list holds more (walker, spin) items than its 2048
    workgroups, which then take several items each; the batch path
    re-inverts with one workgroup per item.  Both give the same chains."""
    s = vmc.System.heg(14, 1.0, 1.0, precision=1, reinvert_interval=8)
    W, seed = 4000, 77
    b = vmc.Batch([s], [W], [seed])
    b.init_positions()
    b.sweep(2, adapt=True, block_size=2)
    b.sweep(3, measure=True, block_size=2)
    got = snapshot_one(b)
    b.close()
    ref = alone(vmc, s, W, seed, 0)
    assert_same([got], [ref])
    assert np.all(ref[1][:, vmc.STAT["NMEAS"]] == 3)
    assert ref[1][:, vmc.STAT["REINV"]].sum() > W


def snapshot_one(b):
    m = b.member(0)
    return (m.positions(), m.stats(), m.steps())


def test_unknown_sweep_flags_are_rejected(vmc):
    """Only VMC_SWEEP_MEASURE and VMC_SWEEP_ADAPT are sweep flags; other bits
    (the library's internal ones among them) fail before any launch."""
    sy = systems(vmc)[:2]
    e = vmc.Ensemble(sy[0], 4, 7)
    e.init_positions()
    b = vmc.Batch(sy, [4, 4], [1, 2])
    b.init_positions()
    L = vmc.lib()
    for bit in (4, 1 << 30):
        assert L.vmc_ensemble_sweep(e._h, 1, bit, 1, e.stream) != 0
        assert L.vmc_batch_sweep(b._h, 1, bit, 1, b.stream) != 0
        assert L.vmc_batch_capture(b._h, 1, bit, 1) != 0
    assert e.sweeps_done() == 0
    e.sweep(1, measure=True)
    assert e.sweeps_done() == 1
    e.close()
    b.close()


def test_batch_members_refuse_direct_sweeps(vmc):
    b = vmc.Batch(systems(vmc)[:2], [4, 4], [1, 2])
    b.init_positions()
    m = b.member(0)
    with pytest.raises(vmc.VmcError):
        m.sweep(1)
    b.close()


def test_batch_mixes_one_and_two_determinant_members(vmc):
    """Generic-kernel batch in mixed precision with K > 0 holding a polarised
    system (one determinant of 19) and an unpolarised one (two of 19): their
    re-inversions form separate sets by order and determinant count, so each
    member equals its alone run and both determinants' drift stays small."""
    import dataclasses

    pol = dataclasses.replace(vmc.System.reference(19, 6.0, precision=vmc.PREC_MIXED), reinvert_interval=8)
    unpol = vmc.System.heg(38, 2.0, 1.0, precision=vmc.PREC_MIXED, reinvert_interval=8)
    sy, W, seeds = [pol, unpol], [24, 24], [31, 32]
    b = vmc.Batch(sy, W, seeds)
    b.set_kernel(vmc.KERNEL_GENERIC)
    b.init_positions()
    b.sweep(2, adapt=True, block_size=2)
    b.sweep(3, measure=True, block_size=2)
    got = [(b.member(i).positions(), b.member(i).stats(), b.member(i).steps()) for i in range(2)]
    drift = [b.member(i).drift() for i in range(2)]
    b.close()
    for i, s in enumerate(sy):
        ref = alone(vmc, s, W[i], seeds[i], vmc.KERNEL_GENERIC)
        assert_same([got[i]], [ref])
        assert np.all(ref[1][:, vmc.STAT["REINV"]] >= 1), i
        assert np.all(drift[i] < 1e-3), (i, drift[i].max())
