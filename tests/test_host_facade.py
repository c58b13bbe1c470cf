"""The C++23 facade's host-only parts against the reference's own classes.

The facade (variational-monte-carlo_amd/host/) restores the reference's
Config, OutputWriter and BlockingAnalysis interfaces above the C ABI.  Its
self-test binary drives them without a device; the reference's classes,
compiled from its sources into oracle/_ref, are driven on the same inputs
through the checker shim, and the results must be identical: the same parsed
fields and derived counts, the same exception kinds, and byte-identical JSON
and binary output (src/output_writer/output_writer.cpp; render.py reads the
binary form).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SELFTEST = os.path.join(ROOT, "variational-monte-carlo_amd", "bin", "vmc_host_selftest")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def selftest(*args):
    assert os.path.exists(SELFTEST), "host facade not built (make -C variational-monte-carlo_amd host)"
    return subprocess.run([SELFTEST, *map(str, args)], capture_output=True, text=True, timeout=60)


def parse_fields(stdout):
    out = {}
    for line in stdout.splitlines():
        if "=" in line and not line.startswith("["):
            k, v = line.split("=", 1)
            out[k] = float(v)
    return out


CONFIGS = {
    "reference_file": open(os.path.join(GOLDEN, "reference_config.cfg")).read(),
    "defaults_and_comments": "# nothing but a comment\n\n   \n",
    "spaces_tabs_unknown": "  Num_Particles\t=  27  # shell\nBox_Length=12.5\nMaster_Seed = 18446744073709551615\n"
                           "Unknown_Key = 3\nJastrow_B = 0.75\nBlock_Size = 7\nWarmup_Sweeps = 0\n",
    "no_equals_and_empty": "Num_Particles 19\nBox_Length =\nJastrow_A = 0.3\n",
    "duplicate_last_wins": "Num_Particles = 7\nNum_Particles = 19\n",
    "bad_particles": "Num_Particles = 0\n",
    "bad_box": "Box_Length = -1\n",
    "bad_block": "Block_Size = 0\n",
    "bad_measure": "Measure_Sweeps = 0\n",
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_config_matches_reference(oracle, tmp_path, name):
    if not oracle.ref_available():
        pytest.skip("reference checker library not built (oracle/_ref)")
    path = tmp_path / "config.cfg"
    path.write_text(CONFIGS[name])
    want = oracle.ref_config(path)
    got = selftest("config", path)
    if want == "invalid_argument":
        assert got.returncode == 3 and "exception invalid_argument" in got.stdout
        return
    assert got.returncode == 0, got.stdout + got.stderr
    fields = parse_fields(got.stdout)
    for k, v in want.items():
        if k == "master_seed":
            # 64-bit seeds: compare through the text form, not a double.
            line = [x for x in got.stdout.splitlines() if x.startswith("master_seed=")][0]
            assert float(line.split("=")[1]) == v
        else:
            assert fields[k] == v, (k, fields[k], v)
    assert ("unknown keys ignored" in got.stdout) == ("Unknown_Key" in CONFIGS[name])


def test_gpu_keys_parse_and_validate(tmp_path):
    path = tmp_path / "c.cfg"
    path.write_text("Num_Particles = 54\nSpin = Unpolarised\nEwald = heg\nMove = gaussian\n"
                    "Precision = mixed\nReinvert_Interval = 64\nNum_Walkers = 4096\n")
    f = parse_fields(selftest("config", path).stdout)
    assert (f["unpolarised"], f["heg_ewald"], f["gaussian_moves"], f["mixed_precision"]) == (1, 1, 1, 1)
    assert f["reinvert_interval"] == 64 and f["walkers"] == 4096
    path.write_text("Num_Particles = 7\nSpin = unpolarised\n")
    r = selftest("config", path)
    assert r.returncode == 3 and "invalid_argument" in r.stdout
    path.write_text("Precision = half\n")
    r = selftest("config", path)
    assert r.returncode == 3 and "invalid_argument" in r.stdout


@pytest.mark.parametrize("fmt", ["json", "bin", "csv"])
def test_writers_byte_identical_to_reference(oracle, tmp_path, fmt):
    if not oracle.ref_available():
        pytest.skip("reference checker library not built (oracle/_ref)")
    ours, theirs = tmp_path / f"ours.{fmt}", tmp_path / f"ref.{fmt}"
    r = selftest("write", fmt, ours)
    err = oracle.ref_write_fixed(fmt, theirs)
    if fmt == "csv":
        assert err == "logic_error"
        assert r.returncode == 3 and "exception logic_error" in r.stdout
        return
    assert err is None and r.returncode == 0
    assert ours.read_bytes() == theirs.read_bytes()


def test_binary_layout_is_the_render_format(tmp_path):
    """Header <QdQ (N, L, measure_steps), then per frame E_L, running mean,
    SE or 0, acceptance and 3N f64 positions (render.py's reader)."""
    out = tmp_path / "o.bin"
    assert selftest("write", "bin", out).returncode == 0
    b = out.read_bytes()
    assert struct.unpack_from("<QdQ", b, 0) == (2, 3.25, 4)
    frame = struct.Struct("<4d6d")
    f1 = frame.unpack_from(b, 24)
    f2 = frame.unpack_from(b, 24 + frame.size)
    assert f1[:4] == (-0.1, -0.1, 0.0, 1.0) and f2[:4] == (0.7, 0.3, 0.125, 0.5)
    assert f2[4:] == (0.15, 0.2, 0.3, 1.0 / 3.0, 2.5, 3.0)
    assert len(b) == 24 + 2 * frame.size


def test_blocking_matches_reference_closed_forms(oracle):
    """Blocks of 2 over 1,3,5,7 -> means 2 and 6 -> mean 4, SE 2
    (tests/test_blocking_analysis.cpp:27-51); the reference's class agrees."""
    r = selftest("blocking", 2, 1, 3, 5, 7)
    f = parse_fields(r.stdout)
    assert f["ready"] == 1 and f["mean"] == 4.0 and f["se"] == 2.0
    x = np.array([1.0, 3.0, 5.0, 7.0, 2.0, 9.0, -4.0, 0.5, 11.0])
    r = selftest("blocking", 3, *x)
    f = parse_fields(r.stdout)
    if oracle.ref_available():
        m, se = oracle.ref_blocking(3, x)
        assert (f["mean"], f["se"]) == (m, se)
    r = selftest("blocking", 5, 1, 2, 3)
    assert r.returncode == 3 and "runtime_error" in r.stdout


@pytest.mark.parametrize("n,L,seed,threads", [(7, 4.0, 5, 2), (19, 12.0, 9, 3), (27, 8.0, 3, 4)])
def test_optimizer_grid_and_choice_match_reference(oracle, tmp_path, n, L, seed, threads):
    """The facade's JastrowOptimizer scans the reference's b grid with the
    reference's sweep counts, and its selection rule (lowest energy unless a
    boundary artifact) picks what the reference's optimize() picks from the
    reference's own energies at those points."""
    if not oracle.ref_available():
        pytest.skip("reference checker library not built (oracle/_ref)")
    b_ref, e_ref, opt_ref = oracle.ref_optimizer(n, L, 0.25, seed, threads)
    cfg = tmp_path / "c.cfg"
    cfg.write_text(f"Num_Threads = {threads}\nNum_Particles = {n}\nBox_Length = {L}\nMaster_Seed = {seed}\n")
    r = selftest("scan", cfg)
    assert r.returncode == 0, r.stdout
    grid = [float(x.split("=")[1]) for x in r.stdout.splitlines() if x.startswith("b=")]
    np.testing.assert_allclose(grid, b_ref, rtol=1e-15, atol=0)
    sweeps = [int(x.split("=")[1]) for x in r.stdout.splitlines() if x.startswith("sweeps=")][0]
    assert sweeps == max(50, 500 // max(n // 7, 1))
    args = [b_ref[0]] + [v for pair in zip(b_ref, e_ref) for v in pair]
    r = selftest("select", *[repr(float(a)) for a in args])
    chosen = float([x for x in r.stdout.splitlines() if x.startswith("b=")][0].split("=")[1])
    assert chosen == opt_ref


def test_optimizer_boundary_rule_cases():
    """Boundary minimum far below a flat interior is an artifact; a boundary
    minimum within twice the interior spread is kept."""
    pick = lambda *a: float(selftest("select", *a).stdout.split("b=")[1].split()[0])
    assert pick(0.5, 0.5, -1.0, 1.0, -0.5, 2.0, -0.51, 3.0, -0.49) == 2.0
    assert pick(0.5, 0.5, -0.52, 1.0, -0.5, 2.0, -0.51, 3.0, -0.49) == 0.5
    assert pick(0.5, 0.5, -0.4, 1.0, -0.5, 2.0, -0.51, 3.0, -0.49) == 2.0
