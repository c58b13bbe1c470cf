"""The per-move objects (vmc_slater, vmc_pade, vmc_ewald) and the facade
classes over them (SlaterPlaneWave, JastrowPade, EnergyTracker, WaveFunction)
on the GPU.

- The reference's unit tests of those classes, restated against the facade
  (host/tests/reference_unit_cases.cpp), all pass.
- A chain of given moves through the reference's per-move call sequence
  (Simulation::metropolis_step, then the measured step's derivative and energy
  calls; simulation.cpp:80-156,212-222) gives, move by move, the ratio, delta J
  and local energy of the reference's own classes (built from their sources as
  a checker, oracle/_ref), and the same inverse and derivatives at the end.
- The C ABI rejects bad arguments and reports a singular determinant as -inf.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "variational-monte-carlo_amd", "bin")
_dp = C.POINTER(C.c_double)


def test_reference_unit_cases_on_the_facade():
    r = subprocess.run([os.path.join(BIN, "vmc_reference_unit_cases")], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 of" in r.stdout


def _chain_inputs(n, L, M, seed, p_accept=0.6, step=0.5):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, L, size=(n, 3))
    cur = pos.copy()
    moved = np.zeros(M, dtype=np.int32)
    newpos = np.zeros((M, 3))
    accept = np.zeros(M, dtype=np.int32)
    for m in range(M):
        k = int(rng.integers(n))
        q = np.mod(cur[k] + rng.uniform(-step, step, 3), L)
        moved[m], newpos[m], accept[m] = k, q, rng.random() < p_accept
        if accept[m]:
            cur[k] = q
    return pos, moved, newpos, accept


@pytest.mark.parametrize("n,L,a,b,M", [(7, 6.0, 0.25, 1.0, 620), (27, 8.0, 0.3, 0.9, 160),
                                       (54, 9.7, 0.25, 1.1, 80)])
def test_move_chain_matches_the_reference_classes(oracle, tmp_path, n, L, a, b, M):
    """620 moves at n = 7 pass the Jastrow cache's 500-call full refresh."""
    if not oracle.ref_available():
        pytest.skip("reference checker library not built (oracle/_ref)")
    pos, moved, newpos, accept = _chain_inputs(n, L, M, seed=100 + n, step=0.4 * L / n ** (1 / 3))
    lines = [f"{n} {L!r} {a!r} {b!r} {M}"]
    lines += [" ".join(repr(float(v)) for v in r) for r in pos]
    lines += [f"{int(k)} {float(q[0])!r} {float(q[1])!r} {float(q[2])!r} {int(c)}"
              for k, q, c in zip(moved, newpos, accept)]
    f = tmp_path / "chain.txt"
    f.write_text("\n".join(lines) + "\n")
    print(f"chain n={n} M={M}", flush=True)
    r = subprocess.run([os.path.join(BIN, "vmc_host_selftest"), "chain", str(f)], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.splitlines()
    mv = np.array([[float(v) for v in x.split()[2:]] for x in out if x.startswith("move ")])
    inv = np.array([[float(v) for v in x.split()[1:]] for x in out if x.startswith("inv ")])
    gl = np.array([[float(v) for v in x.split()[1:]] for x in out if x.startswith("gl ")])
    ratio, dj, e, r_inv, r_grad, r_lap = oracle.ref_move_chain(pos, L, a, b, moved, newpos, accept)
    assert mv.shape == (M, 3)
    np.testing.assert_allclose(mv[:, 0], ratio, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(mv[:, 1], dj, rtol=0, atol=1e-11)
    scale = max(1.0, np.abs(e).max())
    np.testing.assert_allclose(mv[:, 2], e, rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(inv, r_inv, rtol=0, atol=1e-9 * np.abs(r_inv).max())
    np.testing.assert_allclose(gl[:, :3], r_grad, rtol=0, atol=1e-9 * max(1.0, np.abs(r_grad).max()))
    np.testing.assert_allclose(gl[:, 3], r_lap, rtol=0, atol=1e-8 * max(1.0, np.abs(r_lap).max()))
    print(f"n={n}: {int(accept.sum())} of {M} accepted; worst E_L difference "
          f"{np.abs(mv[:, 2] - e).max():.1e}", flush=True)


def _bind(vmc):
    L = vmc.lib()
    L.vmc_slater_create.argtypes = [C.c_int32, C.c_double, C.POINTER(C.c_void_p)]
    L.vmc_slater_destroy.argtypes = [C.c_void_p]
    L.vmc_slater_log_abs_det.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.vmc_slater_update_trig.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double]
    L.vmc_slater_read.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.vmc_slater_shape.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]
    L.vmc_pade_create.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_void_p)]
    L.vmc_pade_destroy.argtypes = [C.c_void_p]
    L.vmc_pade_delta_value.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, C.c_int32, C.c_double,
                                       C.c_double, C.c_double, _dp]
    L.vmc_ewald_create.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_void_p)]
    L.vmc_ewald_destroy.argtypes = [C.c_void_p]
    L.vmc_ewald_init_structure_factors.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp]
    L.vmc_ewald_init_reciprocal.argtypes = [C.c_void_p]
    L.vmc_ewald_init_real.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp]
    L.vmc_ewald_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32), _dp, _dp]
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def test_per_move_c_abi_errors_and_singular(vmc):
    L = _bind(vmc)
    h = C.c_void_p()
    assert L.vmc_slater_create(0, 5.0, C.byref(h)) == vmc.ERR_INVALID
    assert L.vmc_slater_create(3, -1.0, C.byref(h)) == vmc.ERR_INVALID
    assert L.vmc_slater_create(1025, 5.0, C.byref(h)) == vmc.ERR_INVALID
    assert L.vmc_slater_create(3, 10.0, C.byref(h)) == 0
    x, y, z = np.array([1.25, 1.25, 4.1]), np.array([2.5, 2.5, 0.9]), np.array([3.75, 3.75, 1.7])
    ld = C.c_double()
    assert L.vmc_slater_log_abs_det(h, _p(x), _p(y), _p(z), C.byref(ld)) == 0
    assert ld.value == -np.inf  # two particles on one spot: duplicate rows
    assert L.vmc_slater_update_trig(h, 3, 0.0, 0.0, 0.0) == vmc.ERR_INVALID
    assert L.vmc_slater_update_trig(h, -1, 0.0, 0.0, 0.0) == vmc.ERR_INVALID
    assert L.vmc_slater_read(h, 9, _p(np.zeros(64))) == vmc.ERR_INVALID
    assert L.vmc_slater_destroy(h) == 0
    j = C.c_void_p()
    assert L.vmc_pade_create(5.0, 0.25, 1.0, C.byref(j)) == 0
    d = C.c_double()
    assert L.vmc_pade_delta_value(j, 3, _p(x), _p(y), _p(z), 3, 0.0, 0.0, 0.0, C.byref(d)) == vmc.ERR_INVALID
    assert L.vmc_pade_delta_value(j, 0, _p(x), _p(y), _p(z), 0, 0.0, 0.0, 0.0, C.byref(d)) == vmc.ERR_INVALID
    assert L.vmc_pade_destroy(j) == 0


def test_ewald_object_matches_the_reference_tracker(vmc, oracle):
    """The device EnergyTracker state after initialisation: G count, V_real,
    V_recip, self and background terms and S_G equal the reference's."""
    if not oracle.ref_available():
        pytest.skip("reference checker library not built (oracle/_ref)")
    L = _bind(vmc)
    rng = np.random.default_rng(4)
    for n, box in ((7, 6.0), (33, 9.1)):
        pos = rng.uniform(0.0, box, size=(n, 3))
        ref = oracle.ref_energy(pos, box)
        t = C.c_void_p()
        assert L.vmc_ewald_create(box, float(n), C.byref(t)) == 0
        xs = [np.ascontiguousarray(pos[:, c]) for c in range(3)]
        assert L.vmc_ewald_init_structure_factors(t, n, *map(_p, xs)) == 0
        assert L.vmc_ewald_init_reciprocal(t) == 0
        assert L.vmc_ewald_init_real(t, n, *map(_p, xs)) == 0
        vr, vk, se, bg = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        ng = C.c_int32()
        sre, sim = np.zeros(4096), np.zeros(4096)
        assert L.vmc_ewald_state(t, C.byref(vr), C.byref(vk), C.byref(se), C.byref(bg), C.byref(ng), _p(sre),
                                 _p(sim)) == 0
        assert L.vmc_ewald_destroy(t) == 0
        assert ng.value == ref["ng"]
        assert abs(vr.value - ref["vreal"]) <= 1e-12 * max(1.0, abs(ref["vreal"]))
        assert abs(vk.value - ref["vrecip"]) <= 1e-12 * max(1.0, abs(ref["vrecip"]))
        assert abs(se.value - ref["self"]) <= 1e-15 * abs(ref["self"])
        assert abs(bg.value - ref["bg"]) <= 1e-15 * abs(ref["bg"])
        np.testing.assert_allclose(sre[:ng.value], ref["s_re"], rtol=0, atol=1e-12 * n)
        np.testing.assert_allclose(sim[:ng.value], ref["s_im"], rtol=0, atol=1e-12 * n)
