"""GPU parity of the kernels the benchmark times, term by term.

The sweep kernels (one-wavefront k_sweep_rr, multi-wavefront k_sweep_rrw,
generic k_sweep) run one sweep on frozen configurations with proposals and
accept uniforms supplied by the test (vmc_ensemble_sweep_given); every
move's determinant ratio, Jastrow change and decision is compared with the
FP64 oracle making the same moves, and so are the inverses, positions and
log|det| after the sweep.  The local-energy kernels (k_measure_rr,
k_measure_rrw, the generic measurement) return kinetic, V_real, V_recip,
E_L and J per walker (vmc_ensemble_measure), compared before and after the
sweep.

Configurations: half uniform in the box, half taken from an equilibrated
FP64 device chain of the same system.  Proposals: Gaussian displacements of
every electron.

FP64 kernels: every quantity within 1e-10 relative (E_L relative to
max(|E_L|, N x 1 mHa)), and the same accept decisions at the same uniforms.
Mixed precision: the device takes the oracle's decisions (uniform 0 forces an
accept, 1e300 a reject) so both make the same moves; per-move values and
the inverse to the budgets below, and the local energy within
1e-6 x max(|E_L|, N x 1 mHa) of the FP64 oracle.
"""
import zlib

import numpy as np
import pytest

from helpers import oracle_given_sweep, oracle_system

pytestmark = pytest.mark.gpu

# (id, system factory, configurations, kernel)
CASES = [
    ("heg14", lambda V, p: V.System.heg(14, 1.0, 1.0, precision=p), 64, "register"),
    ("heg54", lambda V, p: V.System.heg(54, 2.0, 1.0, precision=p), 64, "register"),
    ("ref27", lambda V, p: V.System.reference(27, 12.0, precision=p), 64, "register"),
    ("heg66", lambda V, p: V.System.heg(66, 2.0, 1.0, precision=p), 64, "register"),
    ("heg114", lambda V, p: V.System.heg(114, 5.0, 0.8, precision=p), 256, "register"),
    ("heg162", lambda V, p: V.System.heg(162, 2.0, 1.0, precision=p), 32, "register"),
    ("heg54_generic", lambda V, p: V.System.heg(54, 2.0, 1.0, precision=p), 64, "generic"),
]
# FP64 with the delayed-update kernel (inverse in HBM, rank-8 Woodbury
# windows): forced at sizes the register kernels also run, and chosen by
# itself at 123 and 147 per spin.
FP64_DELAYED = [
    ("heg14_delayed", lambda V, p: V.System.heg(14, 1.0, 1.0, precision=p), 64, "delayed"),
    ("heg54_delayed", lambda V, p: V.System.heg(54, 2.0, 1.0, precision=p), 64, "delayed"),
    ("ref27_delayed", lambda V, p: V.System.reference(27, 12.0, precision=p), 64, "delayed"),
    ("heg114_delayed", lambda V, p: V.System.heg(114, 5.0, 0.8, precision=p), 32, "delayed"),
    ("heg246", lambda V, p: V.System.heg(246, 2.0, 1.0, precision=p), 16, "auto"),
    ("heg294", lambda V, p: V.System.heg(294, 2.0, 1.0, precision=p), 16, "auto"),
]
MIXED_ONLY = [
    ("heg246", lambda V, p: V.System.heg(246, 2.0, 1.0, precision=p), 16, "register"),
    ("heg294", lambda V, p: V.System.heg(294, 2.0, 1.0, precision=p), 16, "register"),
]

# Mixed precision per-move and inverse budgets against the FP64 oracle:
# the ratio relative to its dot product's rounding scale sum_m |phi_m x_m|
# (the first move of each spin, from the freshly built inverse, and the worst
# move of the sweep, whose inverse carries the FP32 Sherman-Morrison drift of
# the moves before it), the Jastrow change relative to max(|J|, 1), the
# inverse after one sweep of FP32 updates (from an FP64-built start, about
# half the moves accepted) relative to its largest element, and log|det|
# tracked through the FP32 ratios relative to max(|log|det||, 1).  The drift
# of an FP32 inverse grows by about 1/|R| on each accepted move near a node;
# the local energy does not see it, being evaluated in FP64 from an inverse
# refined to FP64 accuracy.
MIXED_MOVE_BUDGET = dict(ratio_first=1e-5, ratio=2e-3, delta_j=1e-6, inverse=1e-3, logdet=1e-4)
E_L_REL = dict(fp64=1e-10, mixed=1e-6)


def frozen_configs(V, s, n_configs, seed):
    """Half uniform configurations, half from 40 adaptive sweeps of an FP64
    device chain of the same system."""
    rng = np.random.default_rng(seed)
    n_uni = n_configs // 2
    uni = rng.uniform(0.0, s.box_length, size=(n_uni, s.N, 3))
    f64 = V.System(**{**s.__dict__, "precision": V.PREC_FP64, "reinvert_interval": 0})
    ens = V.Ensemble(f64, n_configs - n_uni, seed + 1, step0=0.3 * s.box_length / s.N ** (1 / 3))
    ens.init_positions()
    ens.sweep(40, adapt=True)
    eq = ens.positions()
    ens.close()
    return np.concatenate([uni, eq])


def proposals(s, pos, seed):
    rng = np.random.default_rng(seed)
    step = 0.35 * s.box_length / s.N ** (1 / 3)
    q = pos + rng.normal(0.0, step, pos.shape)
    q -= s.box_length * np.floor(q / s.box_length)
    return np.clip(q, 0.0, np.nextafter(s.box_length, 0.0))


def run(V, O, factory, precision, n_configs, kernel, seed):
    s = factory(V, precision)
    pos = frozen_configs(V, s, n_configs, seed)
    prop = proposals(s, pos, seed + 7)
    u = np.random.default_rng(seed + 11).uniform(0.0, 1.0, (n_configs, s.N))
    osys = oracle_system(O, s)
    ref = oracle_given_sweep(O, osys, pos, prop, u)
    ens = V.Ensemble(s, n_configs, seed)
    if kernel == "generic":
        ens.set_kernel(V.KERNEL_GENERIC)
    elif kernel == "delayed":
        ens.set_kernel(V.KERNEL_DELAYED)
    expect = {"generic": V.KERNEL_GENERIC, "delayed": V.KERNEL_DELAYED, "auto": V.KERNEL_DELAYED}
    assert ens.kernel() == expect.get(kernel, V.KERNEL_REGISTER)
    ens.set_positions(pos)
    before = ens.measure()
    if precision == V.PREC_FP64:
        trace = ens.sweep_given(prop, u)
    else:
        acc = np.array([r["trace"][:, 2] for r in ref])
        trace = ens.sweep_given(prop, np.where(acc > 0, 0.0, 1e300))
    inv = ens.inverse()
    got_pos = ens.positions()
    after = ens.measure()
    ens.close()
    return s, pos, ref, before, trace, after, inv, got_pos


def energy_errors(V, O, s, pos_list, got, ref_parts, tag):
    """Relative errors of the measured parts against the oracle's."""
    N = s.N
    kin = np.array([r["kinetic"] for r in ref_parts])
    vre = np.array([r["vreal"] for r in ref_parts])
    vrc = np.array([r["vrecip"] for r in ref_parts])
    el = np.array([r["e_local"] for r in ref_parts])
    J = np.array([r["jastrow"] for r in ref_parts])
    floor = 1e-3 * N
    return {
        f"{tag}kinetic": np.abs(got["KINETIC"] - kin) / np.maximum(np.abs(kin), floor),
        f"{tag}vreal": np.abs(got["VREAL"] - vre) / np.maximum(np.abs(vre), floor),
        f"{tag}vrecip": np.abs(got["VRECIP"] - vrc) / np.maximum(np.abs(vrc), floor),
        f"{tag}e_local": np.abs(got["ELOCAL"] - el) / np.maximum(np.abs(el), floor),
        f"{tag}jastrow": np.abs(got["JASTROW"] - J) / np.maximum(np.abs(J), 1.0),
    }


def oracle_parts(O, s, pos):
    osys = oracle_system(O, s)
    out = []
    for c in range(pos.shape[0]):
        w = O.Walker(osys, pos[c])
        e, t, _ = w.local_energy()
        out.append(dict(kinetic=t, vreal=w.vreal(), vrecip=w.vrecip(), e_local=e, jastrow=w.jastrow(),
                        logdet=[w.logdet(d) for d in range(2 if s.n_dn else 1)]))
    return out


def check(V, O, s, precision, pos, ref, before, trace, after, inv, got_pos):
    ndet = 2 if s.n_dn else 1
    n = [s.n_up, s.n_dn]
    errs = {}
    # Before the sweep: the measurement kernel on the set configurations.
    ref0 = oracle_parts(O, s, pos)
    errs.update(energy_errors(V, O, s, pos, before, ref0, "before_"))
    for d in range(ndet):
        ld = np.array([r["logdet"][d] for r in ref0])
        errs[f"before_logdet{d}"] = np.abs(before[f"LOGDET{d}"] - ld) / np.maximum(np.abs(ld), 1.0)
    # The moves.
    tr_ref = np.array([r["trace"] for r in ref])
    J0 = np.array([r["jastrow"] for r in ref0])
    # The ratio against the rounding scale of its dot product,
    # sum_m |phi_m| |invT[p][m]| (>= |R|; near a node it is the one that counts).
    rscale = np.array([r["ratio_scale"] for r in ref])
    rerr = np.abs(trace[:, :, 0] - tr_ref[:, :, 0]) / rscale
    errs["ratio"] = rerr.max(axis=1)
    # The first move of each spin reads the inverse as built from the
    # configuration (FP64, stored in the working precision): its ratio shows
    # the dot product's own rounding, before any update has drifted it.
    first = [0] + ([s.n_up] if s.n_dn else [])
    errs["ratio_first"] = rerr[:, first].max(axis=1)
    errs["delta_j"] = (np.abs(trace[:, :, 1] - tr_ref[:, :, 1]).max(axis=1) /
                       np.maximum(np.abs(J0), 1.0))
    # After the sweep: state and the measurement.
    assert np.array_equal(trace[:, :, 2], tr_ref[:, :, 2]), "accept decisions differ from the oracle's"
    assert np.array_equal(got_pos, np.array([r["positions"] for r in ref])), "positions after the sweep"
    off = 0
    for d in range(ndet):
        g = inv[:, off:off + n[d] * n[d]].reshape(-1, n[d], n[d])
        off += n[d] * n[d]
        w = np.array([r["inverse"][d] for r in ref])
        errs[f"inverse{d}"] = np.abs(g - w).max(axis=(1, 2)) / np.abs(w).max(axis=(1, 2))
        ld = np.array([r["logdet"][d] for r in ref])
        errs[f"logdet{d}"] = np.abs(after[f"LOGDET{d}"] - ld) / np.maximum(np.abs(ld), 1.0)
    errs.update(energy_errors(V, O, s, got_pos, after, ref, "after_"))
    return {k: float(np.max(v)) for k, v in errs.items()}, errs


def budget_for(V, s, precision):
    """Tolerance of each compared quantity."""
    if precision == V.PREC_FP64:
        return lambda k: 1e-10
    mb = MIXED_MOVE_BUDGET

    def b(k):
        if k.startswith("before_logdet"):
            return 1e-9  # FP64 build of the set configurations
        for key in ("ratio_first", "ratio", "delta_j", "inverse", "logdet"):
            if k.startswith(key):
                return mb[key]
        return E_L_REL["mixed"]  # kinetic, V_real, V_recip, E_L, J
    return b


def seed_of(name):
    return zlib.crc32(name.encode()) % 10000


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,factory,n_configs,kernel", CASES, ids=[c[0] for c in CASES])
def test_timed_kernels_match_oracle(vmc, oracle, name, factory, n_configs, kernel, precision):
    prec = vmc.PREC_FP64 if precision == "fp64" else vmc.PREC_MIXED
    s, pos, ref, before, trace, after, inv, got_pos = run(vmc, oracle, factory, prec, n_configs, kernel,
                                                          seed=seed_of(name))
    worst, _ = check(vmc, oracle, s, prec, pos, ref, before, trace, after, inv, got_pos)
    print(name, precision, " ".join(f"{k}={v:.1e}" for k, v in worst.items()), flush=True)
    b = budget_for(vmc, s, prec)
    bad = {k: v for k, v in worst.items() if not v <= b(k)}
    assert not bad, f"{name} {precision}: beyond budget: {bad}"


@pytest.mark.parametrize("name,factory,n_configs,kernel", MIXED_ONLY, ids=[c[0] for c in MIXED_ONLY])
def test_timed_mixed_large_match_oracle(vmc, oracle, name, factory, n_configs, kernel):
    s, pos, ref, before, trace, after, inv, got_pos = run(vmc, oracle, factory, vmc.PREC_MIXED, n_configs,
                                                          kernel, seed=seed_of(name))
    worst, _ = check(vmc, oracle, s, vmc.PREC_MIXED, pos, ref, before, trace, after, inv, got_pos)
    print(name, "mixed", " ".join(f"{k}={v:.1e}" for k, v in worst.items()), flush=True)
    b = budget_for(vmc, s, vmc.PREC_MIXED)
    bad = {k: v for k, v in worst.items() if not v <= b(k)}
    assert not bad, f"{name} mixed: beyond budget: {bad}"


@pytest.mark.parametrize("name,factory,n_configs,kernel", FP64_DELAYED, ids=[c[0] for c in FP64_DELAYED])
def test_timed_delayed_update_match_oracle(vmc, oracle, name, factory, n_configs, kernel):
    """The FP64 delayed-update sweep: every move's ratio and Jastrow change,
    the decisions, the inverse after the sweep's Woodbury flushes and the
    local energy, each within 1e-10 of the oracle's rank-1 updates."""
    s, pos, ref, before, trace, after, inv, got_pos = run(vmc, oracle, factory, vmc.PREC_FP64, n_configs,
                                                          kernel, seed=seed_of(name))
    worst, _ = check(vmc, oracle, s, vmc.PREC_FP64, pos, ref, before, trace, after, inv, got_pos)
    print(name, "fp64", " ".join(f"{k}={v:.1e}" for k, v in worst.items()), flush=True)
    bad = {k: v for k, v in worst.items() if not v <= 1e-10}
    assert not bad, f"{name} fp64: beyond budget: {bad}"
