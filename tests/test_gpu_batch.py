"""Uneven batches: many systems' walker ensembles swept by one persistent
launch per sweep, optionally captured into a HIP graph.

A walker's chain depends only on its system, seed, walker index and sweep
index, and the batch kernel runs the generic kernel's per-walker routine, so
every system of a batch must come out exactly as when it runs alone with the
generic kernel -- positions, accept counts, step sizes and local energies.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # (N, r_s, b, walkers, seed)
    (14, 1.0, 1.0, 17, 101),
    (38, 3.0, 1.5, 9, 202),
    (66, 2.0, 0.8, 5, 303),
    (54, 5.0, 0.5, 12, 404),
]


def systems(V, precision=0, K=0):
    return [V.System.heg(N, rs, b, precision=precision, reinvert_interval=K) for N, rs, b, _, _ in SPECS]


def run_alone(V, sy, W, seed, sweeps, measure, adapt):
    e = V.Ensemble(sy, W, seed)
    e.set_kernel(V.KERNEL_GENERIC)
    e.init_positions()
    for _ in range(sweeps):
        e.sweep(1, measure=measure, adapt=adapt, block_size=2)
    out = (e.positions(), e.stats(), e.steps())
    e.close()
    return out


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


@pytest.mark.parametrize("precision,K", [(0, 0), (1, 32)], ids=["fp64", "mixed_K32"])
def test_batch_equals_each_system_alone(vmc, precision, K):
    sy = systems(vmc, precision, K)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    b = vmc.Batch(sy, W, seeds)
    info = b.info()
    print("batch", info, flush=True)
    assert info["total_walkers"] == sum(W) and 1 <= info["grid"] <= sum(W)
    b.init_positions()
    b.sweep(2, adapt=True, block_size=2)
    b.sweep(3, measure=True, block_size=2)
    got = snapshot(b)
    b.close()
    for i, (spec, s) in enumerate(zip(SPECS, sy)):
        e = vmc.Ensemble(s, spec[3], spec[4])
        e.set_kernel(vmc.KERNEL_GENERIC)
        e.init_positions()
        e.sweep(2, adapt=True, block_size=2)
        e.sweep(3, measure=True, block_size=2)
        alone = (e.positions(), e.stats(), e.steps())
        e.close()
        assert_same([got[i]], [alone])
        assert np.all(alone[1][:, vmc.STAT["NMEAS"]] == 3)


def test_batch_inverse_placement_hbm_is_identical(vmc):
    sy = systems(vmc)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    outs = []
    for placement in (None, [2] * len(SPECS)):
        b = vmc.Batch(sy, W, seeds, placement=placement)
        b.init_positions()
        b.sweep(3, measure=True, adapt=True)
        outs.append(snapshot(b))
        b.close()
    assert_same(outs[0], outs[1])


def test_batch_graph_replays_continue_the_chains(vmc):
    """A captured graph of 2 sweeps, replayed 3 times with eager reads
    between the replays, equals 6 eager batch sweeps after every replay."""
    import torch

    sy = systems(vmc, precision=1, K=64)
    W = [w for *_, w, _ in SPECS]
    seeds = [s for *_, s in SPECS]
    stream = torch.cuda.Stream()
    eager = vmc.Batch(sy, W, seeds, stream=stream.cuda_stream)
    graph = vmc.Batch(sy, W, seeds, stream=stream.cuda_stream)
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


def test_batch_members_refuse_direct_sweeps(vmc):
    b = vmc.Batch(systems(vmc)[:2], [4, 4], [1, 2])
    b.init_positions()
    m = b.member(0)
    with pytest.raises(vmc.VmcError):
        m.sweep(1)
    b.close()
