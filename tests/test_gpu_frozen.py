"""GPU parity of the per-move and energy kernels on frozen configurations.

Every quantity vmc_frozen_eval returns is compared with the FP64 oracle
(oracle/vmc_oracle.c, itself pinned to the reference) on seeded uniform
configurations, each with one proposed move.  FP64 path: within 1e-10
relative (the tolerance the target states for the GPU FP64 mode).  Mixed
path: the error budget is stated per quantity below.
"""
import numpy as np
import pytest

from helpers import oracle_frozen, random_configs

pytestmark = pytest.mark.gpu

# (name, system factory) pairs: the reference's polarised physics and the
# unpolarised two-determinant HEG of the benchmark configurations.
CASES = [
    ("ref_N7", lambda V, p: V.System.reference(7, 6.0, precision=p)),
    ("ref_N19", lambda V, p: V.System.reference(19, 9.0, a=0.3, b=0.9, precision=p)),
    ("ref_N27", lambda V, p: V.System.reference(27, 12.0, precision=p)),
    ("heg_N14_rs1", lambda V, p: V.System.heg(14, 1.0, 1.0, precision=p)),
    ("heg_N54_rs2", lambda V, p: V.System.heg(54, 2.0, 1.0, precision=p)),
]


def oracle_system(O, s):
    return O.System(s.n_up, s.n_dn, s.box_length, s.a_same, s.a_anti, s.b, s.alpha,
                    s.recip_nmax2, s.real_cutoff, s.erfc_kind)


def run_case(V, O, factory, precision, n_configs, seed):
    s = factory(V, precision)
    rng = np.random.default_rng(seed)
    pos, moved, newpos = random_configs(rng, n_configs, s.N, s.box_length)
    got = V.frozen_eval(s, pos, moved, newpos)
    ref = oracle_frozen(O, oracle_system(O, s), pos, moved, newpos)
    return s, got, ref


def gather(ref, key):
    return np.array([r[key] for r in ref])


def scaled_err(got, want, scale):
    """|got - want| / scale, scale a per-configuration magnitude."""
    return np.abs(np.asarray(got) - np.asarray(want)) / np.asarray(scale)


def check_fp64(s, got, ref, tol=1e-10):
    assert not np.any(got["SINGULAR"]), "unexpected singular configuration"
    ndet = 2 if s.n_dn else 1
    errs = {}
    for d in range(ndet):
        errs[f"logdet{d}"] = scaled_err(got[f"LOGDET{d}"], [r["logdet"][d] for r in ref],
                                        np.maximum(np.abs([r["logdet"][d] for r in ref]), 1.0))
        for tag in ("inv_before", "inv_after"):
            g = got.inv_before[d] if tag == "inv_before" else got.inv_after[d]
            w = np.stack([r[tag][d] for r in ref])
            errs[f"{tag}{d}"] = (np.abs(g - w).max(axis=(1, 2)) /
                                 np.abs(w).max(axis=(1, 2)))
    N = s.N
    # Energies: a term is compared with its own magnitude; E_L, a sum with
    # cancellation, with the magnitude sum of its parts.
    kin, vre, vrc, el = (gather(ref, k) for k in ("kinetic", "vreal", "vrecip", "e_local"))
    escale = np.abs(kin) + np.abs(vre) + np.abs(vrc) + np.abs(el - kin - vre - vrc)
    errs["jastrow"] = scaled_err(got["JASTROW"], gather(ref, "jastrow"),
                                 np.maximum(np.abs(gather(ref, "jastrow")), 1.0))
    errs["kinetic"] = scaled_err(got["KINETIC"], kin, np.maximum(np.abs(kin), 1e-3 * N))
    errs["vreal"] = scaled_err(got["VREAL"], vre, np.maximum(np.abs(vre), 1e-3 * N))
    errs["vrecip"] = scaled_err(got["VRECIP"], vrc, np.maximum(np.abs(vrc), 1e-3 * N))
    errs["e_local"] = scaled_err(got["ELOCAL"], gather(ref, "e_local"), escale)
    errs["ratio"] = scaled_err(got["RATIO"], gather(ref, "ratio"),
                               np.maximum(np.abs(gather(ref, "ratio")), 1e-3))
    errs["delta_j"] = scaled_err(got["DELTA_J"], gather(ref, "delta_j"),
                                 np.maximum(np.abs(gather(ref, "jastrow")), 1.0))
    errs["vreal_inc"] = scaled_err(got["VREAL_INC"], gather(ref, "vreal_inc"),
                                   np.maximum(np.abs(gather(ref, "vreal_inc")), 1e-3 * N))
    errs["vrecip_inc"] = scaled_err(got["VRECIP_INC"], gather(ref, "vrecip_inc"),
                                    np.maximum(np.abs(gather(ref, "vrecip_inc")), 1e-3 * N))
    escale1 = (np.abs(gather(ref, "kinetic_after")) + np.abs(gather(ref, "vreal_inc")) +
               np.abs(gather(ref, "vrecip_inc")) + np.abs(el - kin - vre - vrc))
    errs["e_local_after"] = scaled_err(got["ELOCAL_AFTER"], gather(ref, "e_local_after"), escale1)
    errs["jastrow_after"] = scaled_err(got["JASTROW_AFTER"], gather(ref, "jastrow_after"),
                                       np.maximum(np.abs(gather(ref, "jastrow_after")), 1.0))
    g = got.grad_lap[:, :, :3]
    gw = np.stack([r["grad"] for r in ref])
    errs["grad"] = np.abs(g - gw).max(axis=(1, 2)) / np.maximum(np.abs(gw).max(axis=(1, 2)), 1.0)
    lw = np.stack([r["lap"] for r in ref])
    errs["lap"] = (np.abs(got.grad_lap[:, :, 3] - lw).max(axis=1) /
                   np.maximum(np.abs(lw).max(axis=1), 1.0))
    skw = np.stack([r["s_k"] for r in ref])
    errs["s_k"] = np.abs(got.s_k - skw).max(axis=(1, 2)) / N
    worst = {k: float(np.max(v)) for k, v in errs.items()}
    print(" ".join(f"{k}={v:.2e}" for k, v in worst.items()))
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"FP64 parity beyond {tol}: {bad}"


@pytest.mark.parametrize("name,factory", CASES, ids=[c[0] for c in CASES])
def test_frozen_fp64_matches_oracle(vmc, oracle, name, factory):
    s, got, ref = run_case(vmc, oracle, factory, vmc.PREC_FP64, 64, seed=1234)
    check_fp64(s, got, ref)


# Mixed precision: FP32 orbitals, ratios, Sherman-Morrison and Jastrow
# change per move (FP64 accumulation); the FP64 inverse build, log|det| and
# the local energy and its parts in FP64.  Stated budget against the FP64
# oracle: per-move values relative to their scale (ratio to max(|R|, 1e-2),
# delta J to max(|J|, 1), the inverse to its largest element), the energies
# relative to max(|E|, N x 1 mHa).
MIXED_BUDGET = dict(logdet=1e-9, inv=2e-5, ratio=2e-5, delta_j=1e-6, e_local=1e-6,
                    kinetic=1e-6, vreal=1e-6, vrecip=1e-6)


@pytest.mark.parametrize("name,factory", CASES, ids=[c[0] for c in CASES])
def test_frozen_mixed_error_budget(vmc, oracle, name, factory):
    s, got, ref = run_case(vmc, oracle, factory, vmc.PREC_MIXED, 64, seed=4321)
    N = s.N
    kin, vre, vrc, el = (gather(ref, k) for k in ("kinetic", "vreal", "vrecip", "e_local"))
    errs = dict(
        logdet=np.max(np.abs(got["LOGDET0"] - [r["logdet"][0] for r in ref]) /
                      np.maximum(np.abs([r["logdet"][0] for r in ref]), 1.0)),
        inv=max(float(np.max(np.abs(got.inv_before[d] - np.stack([r["inv_before"][d] for r in ref]))
                             / np.abs(np.stack([r["inv_before"][d] for r in ref])).max()))
                for d in range(len(got.inv_before))),
        ratio=np.max(np.abs(got["RATIO"] - gather(ref, "ratio")) /
                     np.maximum(np.abs(gather(ref, "ratio")), 1e-2)),
        delta_j=np.max(np.abs(got["DELTA_J"] - gather(ref, "delta_j")) /
                       np.maximum(np.abs(gather(ref, "jastrow")), 1.0)),
        e_local=np.max(np.abs(got["ELOCAL"] - el) / np.maximum(np.abs(el), 1e-3 * N)),
        kinetic=np.max(np.abs(got["KINETIC"] - kin) / np.maximum(np.abs(kin), 1e-3 * N)),
        vreal=np.max(np.abs(got["VREAL"] - vre) / np.maximum(np.abs(vre), 1e-3 * N)),
        vrecip=np.max(np.abs(got["VRECIP"] - vrc) / np.maximum(np.abs(vrc), 1e-3 * N)),
    )
    print(" ".join(f"{k}={float(v):.2e}" for k, v in errs.items()))
    bad = {k: float(v) for k, v in errs.items() if not v <= MIXED_BUDGET[k]}
    assert not bad, f"mixed precision beyond budget {MIXED_BUDGET}: {bad}"


def test_n114_mixed_error_report(vmc, oracle):
    """The mixed-precision configuration of N = 114 (57 + 57, |n|^2 <= 5
    shells), r_s = 5, b = 0.8: ratio, log|det| and local energy against the
    FP64 oracle on 256 seeded frozen configurations, to the mixed budget,
    except the ratio: a 57-term FP32 dot against an FP32-rounded inverse,
    held to 5e-5 of max(|R|, 1e-2)."""
    factory = lambda V, p: V.System.heg(114, 5.0, 0.8, precision=p)
    s, got, ref = run_case(vmc, oracle, factory, vmc.PREC_MIXED, 256, seed=114)
    assert not np.any(got["SINGULAR"])
    el = gather(ref, "e_local")
    errs = dict(
        ratio=np.max(np.abs(got["RATIO"] - gather(ref, "ratio")) /
                     np.maximum(np.abs(gather(ref, "ratio")), 1e-2)),
        logdet=max(np.max(np.abs(got[f"LOGDET{d}"] - [r["logdet"][d] for r in ref]) /
                          np.maximum(np.abs([r["logdet"][d] for r in ref]), 1.0)) for d in (0, 1)),
        e_local=np.max(np.abs(got["ELOCAL"] - el) / np.maximum(np.abs(el), 1e-3 * s.N)),
    )
    print(" ".join(f"{k}={float(v):.2e}" for k, v in errs.items()))
    budget = dict(ratio=5e-5, logdet=MIXED_BUDGET["logdet"], e_local=MIXED_BUDGET["e_local"])
    bad = {k: float(v) for k, v in errs.items() if not v <= budget[k]}
    assert not bad, f"N=114 mixed precision beyond budget {budget}: {bad}"
