"""C ABI checks that need no GPU: the library loads, exports every function
include/vmc_hip.h declares, derives the same system sizes as the oracle, and
rejects invalid arguments before touching a device."""
import ctypes as C
import subprocess

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(vmc_nodev):
    V = vmc_nodev
    declared = V.declared_functions()
    assert len(declared) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", V.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    L = V.lib()
    for f in declared:
        assert getattr(L, f) is not None


@pytest.mark.parametrize("n", [1, 7, 14, 38, 54, 66, 114, 162, 246, 294])
def test_system_sizes_match_oracle(vmc_nodev, oracle, n):
    V = vmc_nodev
    s = V.System.heg(n, 2.0, 1.0) if n % 2 == 0 else V.System.reference(n, 7.0)
    (k0, k1), ng, ninv = V.system_sizes(s)
    osys = oracle.System(s.n_up, s.n_dn, s.box_length, s.a_same, s.a_anti, s.b, s.alpha,
                         s.recip_nmax2, s.real_cutoff, s.erfc_kind)
    w = oracle.Walker(osys)
    assert ng == w.ng
    assert k0 == oracle.shell_fill(s.n_up)[0].shape[0]
    if s.n_dn:
        assert k1 == oracle.shell_fill(s.n_dn)[0].shape[0]
    assert ninv == s.n_up ** 2 + s.n_dn ** 2


def test_reference_rule_gives_751_g_vectors(vmc_nodev):
    (_, _), ng, _ = vmc_nodev.system_sizes(vmc_nodev.System.reference(54, 12.0))
    assert ng == 751
    (_, _), ng, _ = vmc_nodev.system_sizes(vmc_nodev.System.heg(54, 2.0, 1.0))
    assert ng == 257


def test_invalid_arguments_rejected_without_device(vmc_nodev):
    V = vmc_nodev
    bad = V.System.reference(0, 5.0)
    with pytest.raises(V.VmcError) as e:
        V.system_sizes(bad)
    assert e.value.code == 1
    neg_box = V.System.reference(7, -1.0)
    h = C.c_void_p()
    d = neg_box.desc()
    assert V.lib().vmc_ensemble_create(C.byref(d), 4, 1, 0.5, C.byref(h)) == 1
    assert b"box_length" in V.lib().vmc_last_error()
    good = V.System.reference(7, 5.0).desc()
    assert V.lib().vmc_ensemble_create(C.byref(good), 0, 1, 0.5, C.byref(h)) == 1
    pos = np.full((1, 7, 3), 1.0)
    with pytest.raises(V.VmcError):
        V.frozen_eval(V.System.reference(7, 5.0), pos, [9], [[1.0, 1.0, 1.0]])
    with pytest.raises(V.VmcError):
        V.frozen_eval(V.System.reference(7, 5.0), pos, [0], [[6.0, 1.0, 1.0]])


def test_footprint_counts_every_per_walker_buffer(vmc_nodev):
    """vmc_ensemble_footprint (the server's admission measure) counts what an
    ensemble allocates per walker: positions, the working-precision inverses,
    the FP64 Gauss-Jordan scratch, statistics, step state, drift, the Jastrow
    row sums, the re-inversion list entry and -- for mixed ensembles -- the
    FP64 inverses of a measured sweep and, above 80 electrons per spin, the
    Newton-Schulz residual (padded order^2 floats)."""
    V = vmc_nodev

    def per_walker(s):
        return V.footprint(s, 2) - V.footprint(s, 1)

    N, n = 54, 27
    s = V.System.heg(N, 2.0, 1.0, precision=V.PREC_MIXED, reinvert_interval=64)
    assert per_walker(s) == (3 * N * 8 + 2 * n * n * 4 + n * n * 8 + 16 * 8 + 8 + 16 + 12 + 8 + 8 * N +
                             2 * n * n * 8)
    s64 = V.System.heg(N, 2.0, 1.0, precision=V.PREC_FP64)
    assert per_walker(s64) - per_walker(s) == 2 * n * n * 4 - 2 * n * n * 8
    for N, p in ((162, 96), (246, 128), (294, 160)):
        k = V.System.heg(N, 2.0, 1.0, precision=V.PREC_MIXED, reinvert_interval=64)
        k64 = V.System.heg(N, 2.0, 1.0, precision=V.PREC_FP64, reinvert_interval=64)
        n = N // 2
        assert per_walker(k) - per_walker(k64) == p * p * 4 + 2 * n * n * 8 - 2 * n * n * 4
