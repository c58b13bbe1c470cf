"""The reference's measurement cadence for one walker: vmc_ensemble_sweep_traced
evaluates the traced walker's local energy after each of its moves
(Simulation::measure, src/simulation/simulation.cpp:203-243).  Each recorded
E_L equals the FP64 oracle's at the recorded positions (1e-10), the recorded
positions change exactly at the accepted moves of the recorded electrons, and
tracing leaves the chain as an untraced sweep on the same kernel leaves it.
The C++ facade streams these per-step frames: the binary output holds
measure_steps frames, as its header states.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import oracle_system, sweep_order

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "variational-monte-carlo_amd", "bin", "vmc_gpu")


@pytest.mark.parametrize("spec", [(14, 1.0, 1.0, 0), (7, 3.0, 1.0, 1), (38, 2.0, 0.8, 0)],
                         ids=["N14_heg", "N7_reference", "N38_heg"])
def test_traced_walker_local_energy_after_every_move(vmc, oracle, spec):
    N, rs, b, reference_rules = spec
    if reference_rules:
        s = vmc.System.reference(N, 6.0, 0.25, b)
    else:
        s = vmc.System.heg(N, rs, b)
    W, seed, traced = 6, 4242, 4
    twin = vmc.Ensemble(s, W, seed, step0=0.4 * s.box_length / 4)
    ens = vmc.Ensemble(s, W, seed, step0=0.4 * s.box_length / 4)
    for e in (ens, twin):
        e.set_kernel(vmc.KERNEL_GENERIC)
        e.init_positions()
        e.sweep(3, adapt=True)
    osys = oracle_system(oracle, s)
    order = sweep_order(s.n_up, s.n_dn)
    for sweep in range(3):
        before = ens.positions()[traced]
        tr = ens.sweep_traced(traced, measure=True, block_size=2)
        twin.sweep(1, measure=True, block_size=2)
        assert list(tr["moved"]) == order
        prev = before
        for t in range(N):
            p = int(tr["moved"][t])
            cur = tr["pos"][t]
            changed = np.any(cur != prev, axis=1)
            if tr["accepted"][t]:
                assert list(np.nonzero(changed)[0]) in ([p], []), (sweep, t)
            else:
                assert not changed.any(), (sweep, t)
            e_or, _, _ = oracle.Walker(osys, cur).local_energy()
            assert abs(tr["e_local"][t] - e_or) <= 1e-10 * max(abs(e_or), 1e-3 * N), (sweep, t, tr["e_local"][t], e_or)
            prev = cur
        np.testing.assert_array_equal(ens.positions()[traced], tr["pos"][-1])
        # Tracing changes nothing of the chain.
        np.testing.assert_array_equal(ens.positions(), twin.positions())
        np.testing.assert_array_equal(ens.stats(), twin.stats())
    assert ens.sweeps_done() == twin.sweeps_done() == 6
    ens.close()
    twin.close()


def test_traced_sweep_needs_fp64(vmc):
    s = vmc.System.heg(14, 1.0, 1.0, precision=vmc.PREC_MIXED, reinvert_interval=64)
    e = vmc.Ensemble(s, 4, 1)
    e.init_positions()
    with pytest.raises(vmc.VmcError) as err:
        e.sweep_traced(0)
    assert err.value.code == vmc.ERR_INVALID
    with pytest.raises(vmc.VmcError):
        vmc.Ensemble(vmc.System.heg(14, 1.0, 1.0), 4, 1).sweep_traced(0)  # no positions yet
    e.close()


def test_cli_streams_a_frame_per_step(tmp_path):
    """Frames = step (the FP64 default): the reference's per-step stream, one
    frame per move of walker 0 with its local energy after that move, the
    running mean of those energies and the acceptance so far."""
    N, L, sweeps = 7, 6.0, 60
    (tmp_path / "output").mkdir()
    (tmp_path / "config.cfg").write_text(f"Num_Walkers = 64\nNum_Particles = {N}\nBox_Length = {L}\n"
                                         f"Warmup_Sweeps = 20\nMeasure_Sweeps = {sweeps}\nBlock_Size = 35\n"
                                         f"Master_Seed = 9\nFrames = step\n")
    r = subprocess.run([CLI, "--no-optimize"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = (tmp_path / "output" / "vmc.bin").read_bytes()
    assert struct.unpack_from("<QdQ", b, 0) == (N, L, N * sweeps)
    frame = 4 + 3 * N
    assert len(b) == 24 + N * sweeps * frame * 8
    frames = np.frombuffer(b, dtype="<f8", offset=24).reshape(N * sweeps, frame)
    running = np.cumsum(frames[:, 0]) / np.arange(1, N * sweeps + 1)
    np.testing.assert_allclose(frames[:, 1], running, rtol=1e-12)
    # Between consecutive frames at most one electron moves.
    pos = frames[:, 4:].reshape(N * sweeps, N, 3)
    moved = np.any(pos[1:] != pos[:-1], axis=2).sum(axis=1)
    assert moved.max() <= 1 and moved.sum() > 0
