"""Python binding of libvmc_hip.so (include/vmc_hip.h) over ctypes.

This is the host-side mirror used by the tests, the benchmark and the smoke
check.  It calls the gfx950 library only; there is no CPU fallback, and every
entry point raises if the library or a gfx950 device is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libvmc_hip.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "vmc_hip.h")

VMC_OK = 0
ERRORS = {1: "invalid argument", 2: "device error", 3: "singular", 4: "out of memory",
          5: "bad state", 6: "cancelled"}
ERR_INVALID, ERR_DEVICE, ERR_SINGULAR, ERR_MEMORY, ERR_STATE, ERR_CANCELLED = 1, 2, 3, 4, 5, 6

PREC_FP64, PREC_MIXED = 0, 1
MOVE_UNIFORM, MOVE_GAUSSIAN = 0, 1
ERFC_AS7126, ERFC_EXACT = 0, 1
SWEEP_MEASURE, SWEEP_ADAPT = 1, 2
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_REGISTER, KERNEL_DELAYED = 0, 1, 2, 3

STAT = dict(NMEAS=0, SUM_E=1, SUM_E2=2, SUM_KIN=3, SUM_POT=4, LAST_E=5, ACCEPTED=6,
            PROPOSED=7, BLK_SUM=8, BLK_IN=9, BLK_N=10, BLK_MEAN=11, BLK_M2=12, LAST_KIN=13,
            LAST_POT=14, REINV=15)
STAT_COUNT = 16
FZ = dict(LOGDET0=0, LOGDET1=1, JASTROW=2, KINETIC=3, VREAL=4, VRECIP=5, ELOCAL=6, RATIO=7,
          DELTA_J=8, VREAL_INC=9, VRECIP_INC=10, ELOCAL_AFTER=11, KINETIC_AFTER=12,
          JASTROW_AFTER=13, SINGULAR=14)
FZ_COUNT = 15
MV = dict(RATIO=0, DELTA_J=1, ACCEPTED=2)
MV_COUNT = 3
PART = dict(KINETIC=0, VREAL=1, VRECIP=2, ELOCAL=3, LOGDET0=4, LOGDET1=5, JASTROW=6)
PART_COUNT = 7
NOALLOC_MAX_WALKERS = 65536  # VMC_NOALLOC_MAX_WALKERS


class VmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vmc error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class SystemDesc(C.Structure):
    """Mirror of vmc_system."""

    _fields_ = [
        ("n_up", C.c_int32),
        ("n_dn", C.c_int32),
        ("box_length", C.c_double),
        ("jastrow_a_same", C.c_double),
        ("jastrow_a_anti", C.c_double),
        ("jastrow_b", C.c_double),
        ("ewald_alpha", C.c_double),
        ("recip_nmax2", C.c_int32),
        ("real_cutoff", C.c_double),
        ("erfc_kind", C.c_int32),
        ("move_kind", C.c_int32),
        ("precision", C.c_int32),
        ("reinvert_interval", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("walker_threads", C.c_int32), ("rr_variant", C.c_int32),
                ("inv_placement", C.c_int32),
                ("reinvert_interval", C.c_int32), ("moves_per_s", C.c_double), ("drift", C.c_double)]


class Request(C.Structure):
    _fields_ = [("system", SystemDesc), ("n_walkers", C.c_int64), ("warmup_sweeps", C.c_int32),
                ("measure_sweeps", C.c_int32), ("block_size", C.c_int32), ("seed", C.c_uint64),
                ("step0", C.c_double)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("walkers", C.c_int64),
                ("mean_energy", C.c_double), ("standard_error", C.c_double),
                ("acceptance_rate", C.c_double), ("moves", C.c_double), ("queued_s", C.c_double),
                ("latency_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def box_length_from_rs(rs, n):
    return float(np.cbrt(4.0 * np.pi * n / 3.0) * rs)


@dataclass
class System:
    """One homogeneous-electron-gas system (vmc_system)."""

    n_up: int
    n_dn: int
    box_length: float
    a_same: float = 0.25
    a_anti: float = 0.5
    b: float = 1.0
    alpha: float = 0.0
    recip_nmax2: int = 0
    real_cutoff: float = 0.0
    erfc_kind: int = ERFC_AS7126
    move_kind: int = MOVE_UNIFORM
    precision: int = PREC_FP64
    reinvert_interval: int = 0

    @property
    def N(self):
        return self.n_up + self.n_dn

    @classmethod
    def reference(cls, n, L, a=0.25, b=1.0, precision=PREC_FP64):
        """The reference's own physics: one polarised determinant, alpha = 6/L,
        tolerance-rule G vectors, all minimum-image pairs, A&S erfc,
        uniform-cube moves."""
        return cls(n, 0, L, a, a, b, 0.0, 0, 0.0, ERFC_AS7126, MOVE_UNIFORM, precision, 0)

    @classmethod
    def heg(cls, n_total, rs, b, precision=PREC_FP64, reinvert_interval=0, a_same=0.25,
            a_anti=0.5, nmax2=25):
        """Unpolarised HEG: two determinants of N/2, alpha = 5/L, rcut = L/2,
        |n|^2 <= nmax2 reciprocal shells, exact erfc, Gaussian moves."""
        L = box_length_from_rs(rs, n_total)
        return cls(n_total // 2, n_total - n_total // 2, L, a_same, a_anti, b, 5.0 / L, nmax2,
                   0.5 * L, ERFC_EXACT, MOVE_GAUSSIAN, precision, reinvert_interval)

    def desc(self):
        return SystemDesc(self.n_up, self.n_dn, self.box_length, self.a_same, self.a_anti, self.b,
                          self.alpha, self.recip_nmax2, self.real_cutoff, self.erfc_kind,
                          self.move_kind, self.precision, self.reinvert_interval)


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_VP = C.c_void_p


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def declared_functions():
    """Names of every function include/vmc_hip.h declares."""
    src = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vmc_\w+)\s*\(", src, re.M)))


def lib():
    """Load libvmc_hip.so; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               f"`make -C {PKG_DIR}` (hipcc, gfx950)")
        # One HIP runtime per process: when PyTorch is present its runtime is
        # loaded first, and the library binds to it (same soname), so torch
        # streams and events and this library share devices and queues.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        sp = C.POINTER(SystemDesc)
        sig = {
            "vmc_last_error": (C.c_char_p, []),
            "vmc_device_check": (C.c_int, [_ip]),
            "vmc_system_sizes": (C.c_int, [sp, _ip, _ip, _ip]),
            "vmc_ensemble_create": (C.c_int, [sp, C.c_int64, C.c_uint64, C.c_double,
                                              C.POINTER(_VP)]),
            "vmc_ensemble_destroy": (C.c_int, [_VP]),
            "vmc_ensemble_init_positions": (C.c_int, [_VP, _VP]),
            "vmc_ensemble_set_positions": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_get_positions": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_sweep": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _VP]),
            "vmc_ensemble_sweep_given": (C.c_int, [_VP, _dp, _dp, _dp, _VP]),
            "vmc_ensemble_measure": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_sweep_traced": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int64, _dp, _ip, _ip,
                                                    _dp, _VP]),
            "vmc_ensemble_measure_noalloc": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_get_inverse": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_reset_stats": (C.c_int, [_VP, _VP]),
            "vmc_ensemble_set_kernel": (C.c_int, [_VP, C.c_int32]),
            "vmc_ensemble_get_kernel": (C.c_int, [_VP, _ip]),
            "vmc_ensemble_set_rr_variant": (C.c_int, [_VP, C.c_int32]),
            "vmc_ensemble_set_walker_groups": (C.c_int, [_VP, C.c_int32]),
            "vmc_ensemble_read_stats": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_read_steps": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_read_drift": (C.c_int, [_VP, _dp, _VP]),
            "vmc_ensemble_set_inverse_placement": (C.c_int, [_VP, C.c_int32]),
            "vmc_ensemble_footprint": (C.c_int, [sp, C.c_int64, C.POINTER(C.c_int64)]),
            "vmc_autotune": (C.c_int, [sp, C.c_int64, C.c_double, C.c_int32, C.POINTER(Tuning), _VP]),
            "vmc_server_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(_VP)]),
            "vmc_server_submit": (C.c_int, [_VP, C.POINTER(Request), C.POINTER(C.c_int64)]),
            "vmc_server_poll": (C.c_int, [_VP, C.POINTER(Result), C.c_int32, _ip, C.c_int32]),
            "vmc_server_cancel": (C.c_int, [_VP, C.c_int64]),
            "vmc_server_destroy": (C.c_int, [_VP]),
            "vmc_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                           _dp, _ip, C.c_int32, C.c_int32, C.POINTER(_VP)]),
            "vmc_ensemble_set_walker_threads": (C.c_int, [_VP, C.c_int32]),
            "vmc_default_walker_threads": (C.c_int, [sp]),
            "vmc_batch_destroy": (C.c_int, [_VP]),
            "vmc_batch_init_positions": (C.c_int, [_VP, _VP]),
            "vmc_batch_sweep": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _VP]),
            "vmc_batch_capture": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32]),
            "vmc_batch_replay": (C.c_int, [_VP, _VP]),
            "vmc_batch_member": (C.c_int, [_VP, C.c_int32, C.POINTER(_VP)]),
            "vmc_batch_info": (C.c_int, [_VP, _ip, C.POINTER(C.c_int64), _ip,
                                         C.POINTER(C.c_int64)]),
            "vmc_batch_set_kernel": (C.c_int, [_VP, C.c_int32, C.c_int32]),
            "vmc_batch_set_overlap": (C.c_int, [_VP, C.c_int32]),
            "vmc_batch_groups": (C.c_int, [_VP, _ip, _ip, _ip, C.c_int32]),
            "vmc_ensemble_set_timing": (C.c_int, [_VP, C.c_int32]),
            "vmc_ensemble_read_timing": (C.c_int, [_VP, _dp, C.POINTER(C.c_int64)]),
            "vmc_ensemble_info": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
            "vmc_frozen_eval": (C.c_int, [sp, C.c_int32, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _VP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != VMC_OK:
        raise VmcError(rc, lib().vmc_last_error().decode())


def device_check():
    n = C.c_int32(0)
    _check(lib().vmc_device_check(C.byref(n)))
    return n.value


def system_sizes(system: System):
    nk = (C.c_int32 * 2)()
    ng, ninv = C.c_int32(), C.c_int32()
    d = system.desc()
    _check(lib().vmc_system_sizes(C.byref(d), nk, C.byref(ng), C.byref(ninv)))
    return (nk[0], nk[1]), ng.value, ninv.value


class Ensemble:
    """Device-resident walker ensemble (vmc_ensemble)."""

    @classmethod
    def _borrowed(cls, handle, system, n_walkers, seed, stream):
        """View of a batch member: reads only, the batch owns and sweeps it."""
        self = cls.__new__(cls)
        self.system, self.W, self.N, self.seed, self.stream = system, int(n_walkers), system.N, seed, stream
        self._desc = system.desc()
        self._h = handle
        self._owned = False
        return self

    def __init__(self, system: System, n_walkers: int, seed: int, step0: float | None = None,
                 stream=None):
        self.system = system
        self.W = int(n_walkers)
        self.N = system.N
        self.seed = seed
        self.stream = stream
        self._h = _VP()
        self._desc = system.desc()
        step = system.box_length / 10.0 if step0 is None else step0
        _check(lib().vmc_ensemble_create(C.byref(self._desc), self.W, seed, step,
                                         C.byref(self._h)))
        self._owned = True

    def close(self):
        if self._h and self._owned:
            lib().vmc_ensemble_destroy(self._h)
        self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_positions(self):
        _check(lib().vmc_ensemble_init_positions(self._h, self.stream))

    def set_positions(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.W, self.N, 3)
        _check(lib().vmc_ensemble_set_positions(self._h, _ptr(p), self.stream))

    def positions(self):
        p = np.zeros((self.W, self.N, 3))
        _check(lib().vmc_ensemble_get_positions(self._h, _ptr(p), self.stream))
        return p

    def sweep(self, n_sweeps=1, measure=False, adapt=False, block_size=1):
        flags = (SWEEP_MEASURE if measure else 0) | (SWEEP_ADAPT if adapt else 0)
        _check(lib().vmc_ensemble_sweep(self._h, n_sweeps, flags, block_size, self.stream))

    def sweep_given(self, proposals, uniforms):
        """One sweep with the caller's moves (vmc_ensemble_sweep_given):
        proposals [W][N][3], uniforms [W][N]; returns the trace [W][N][MV_COUNT]
        (ratio, Jastrow change, accepted) indexed by electron."""
        p = np.ascontiguousarray(proposals, dtype=np.float64).reshape(self.W, self.N, 3)
        u = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(self.W, self.N)
        tr = np.zeros((self.W, self.N, MV_COUNT))
        _check(lib().vmc_ensemble_sweep_given(self._h, _ptr(p), _ptr(u), _ptr(tr), self.stream))
        return tr

    def sweep_traced(self, walker, measure=True, adapt=False, block_size=1, positions=True):
        """One sweep (vmc_ensemble_sweep_traced, generic kernel, FP64) with
        walker `walker`'s local energy after each of its moves: a dict of
        e_local [N], moved [N], accepted [N] and, with positions, pos [N][N][3]
        (the walker after each move), in sweep order."""
        flags = (SWEEP_MEASURE if measure else 0) | (SWEEP_ADAPT if adapt else 0)
        el = np.zeros(self.N)
        mv = np.zeros(self.N, dtype=np.int32)
        acc = np.zeros(self.N, dtype=np.int32)
        pos = np.zeros((self.N, self.N, 3)) if positions else None
        _check(lib().vmc_ensemble_sweep_traced(self._h, flags, block_size, int(walker), _ptr(el),
                                               mv.ctypes.data_as(_ip), acc.ctypes.data_as(_ip),
                                               _ptr(pos) if positions else None, self.stream))
        out = dict(e_local=el, moved=mv, accepted=acc)
        if positions:
            out["pos"] = pos
        return out

    def measure(self):
        """Per-walker local energy parts (vmc_ensemble_measure) as a dict of
        arrays keyed like PART."""
        a = np.zeros((self.W, PART_COUNT))
        _check(lib().vmc_ensemble_measure(self._h, _ptr(a), self.stream))
        return {k: a[:, i].copy() for k, i in PART.items()}

    def measure_noalloc(self):
        """measure() through vmc_ensemble_measure_noalloc: the same parts,
        staged in the library's per-device array instead of an allocation
        (at most NOALLOC_MAX_WALKERS walkers)."""
        a = np.zeros((self.W, PART_COUNT))
        _check(lib().vmc_ensemble_measure_noalloc(self._h, _ptr(a), self.stream))
        return {k: a[:, i].copy() for k, i in PART.items()}

    def inverse(self):
        """Transposed inverses [W][n_up^2 + n_dn^2] as doubles."""
        (_, _), _, ninv = system_sizes(self.system)
        a = np.zeros((self.W, ninv))
        _check(lib().vmc_ensemble_get_inverse(self._h, _ptr(a), self.stream))
        return a

    def set_kernel(self, kernel):
        _check(lib().vmc_ensemble_set_kernel(self._h, kernel))

    def set_walker_threads(self, threads):
        """Generic kernel: 64, 128 or 256 threads (1, 2 or 4 wavefronts) per walker."""
        _check(lib().vmc_ensemble_set_walker_threads(self._h, int(threads)))

    def set_inverse_placement(self, placement):
        """Generic kernel: 0 LDS when it fits, 1 LDS, 2 HBM."""
        _check(lib().vmc_ensemble_set_inverse_placement(self._h, int(placement)))

    def kernel(self):
        k = C.c_int32()
        _check(lib().vmc_ensemble_get_kernel(self._h, C.byref(k)))
        return k.value

    def set_rr_variant(self, variant):
        _check(lib().vmc_ensemble_set_rr_variant(self._h, variant))

    def set_walker_groups(self, groups):
        """Walker groups (1 to 8) of the register kernels' measured sweeps,
        each on a stream of its own; the chains do not depend on it."""
        _check(lib().vmc_ensemble_set_walker_groups(self._h, int(groups)))

    def reset_stats(self):
        _check(lib().vmc_ensemble_reset_stats(self._h, self.stream))

    def stats(self):
        s = np.zeros((self.W, STAT_COUNT))
        _check(lib().vmc_ensemble_read_stats(self._h, _ptr(s), self.stream))
        return s

    def steps(self):
        s = np.zeros(self.W)
        _check(lib().vmc_ensemble_read_steps(self._h, _ptr(s), self.stream))
        return s

    def drift(self):
        """Largest |I - D^-1 D| (max norm) met per walker when its
        mixed-precision inverse came due for the FP64 re-inversion."""
        s = np.zeros(self.W)
        _check(lib().vmc_ensemble_read_drift(self._h, _ptr(s), self.stream))
        return s

    def set_timing(self, on=True):
        """Record HIP events around every kernel launch on the ensemble's stream."""
        _check(lib().vmc_ensemble_set_timing(self._h, 1 if on else 0))

    def timing(self):
        """{'sweep' | 'reinvert' | 'measure': (ms, launches)} since the last read."""
        ms = np.zeros(3)
        n = np.zeros(3, dtype=np.int64)
        _check(lib().vmc_ensemble_read_timing(self._h, _ptr(ms), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(("sweep", "reinvert", "measure"))}

    def sweeps_done(self):
        w, s = C.c_int64(), C.c_uint64()
        _check(lib().vmc_ensemble_info(self._h, C.byref(w), C.byref(s)))
        return s.value


@dataclass
class FrozenResult:
    scalars: np.ndarray
    grad_lap: np.ndarray
    inv_before: list
    inv_after: list
    s_k: np.ndarray

    def __getitem__(self, key):
        return self.scalars[:, FZ[key]]


def frozen_eval(system: System, pos, moved, newpos, stream=None):
    """Frozen-configuration evaluation (vmc_frozen_eval) of C configurations."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Cn = pos.shape[0]
    N = system.N
    pos = pos.reshape(Cn, N, 3)
    moved = np.ascontiguousarray(moved, dtype=np.int32).reshape(Cn)
    newpos = np.ascontiguousarray(newpos, dtype=np.float64).reshape(Cn, 3)
    (_, _), ng, ninv = system_sizes(system)
    sc = np.zeros((Cn, FZ_COUNT))
    gl = np.zeros((Cn, N, 4))
    ib = np.zeros((Cn, ninv))
    ia = np.zeros((Cn, ninv))
    sk = np.zeros((Cn, max(ng, 1), 2))
    d = system.desc()
    _check(lib().vmc_frozen_eval(C.byref(d), Cn, _ptr(pos), _ptr(moved, _ip), _ptr(newpos),
                                 _ptr(sc), _ptr(gl), _ptr(ib), _ptr(ia), _ptr(sk), stream))
    n0, n1 = system.n_up, system.n_dn

    def split(a):
        out = [a[:, : n0 * n0].reshape(Cn, n0, n0)]
        if n1:
            out.append(a[:, n0 * n0:].reshape(Cn, n1, n1))
        return out

    return FrozenResult(sc, gl, split(ib), split(ia), sk[:, :ng])


def blocking_summary(stats):
    """Mean of per-walker blocked means and the pooled standard error
    sqrt(sum SE_w^2) / W, as main.cpp:70-104 aggregates thread summaries."""
    nb = stats[:, STAT["BLK_N"]]
    ok = nb >= 2
    means = np.where(ok, stats[:, STAT["BLK_MEAN"]], stats[:, STAT["SUM_E"]] /
                     np.maximum(stats[:, STAT["NMEAS"]], 1))
    mean = float(np.mean(means))
    if not np.any(ok):
        return mean, None
    var = stats[ok, STAT["BLK_M2"]] / (nb[ok] - 1)
    se_w = np.sqrt(var / nb[ok])
    se = float(math.sqrt(np.sum(se_w ** 2)) / stats.shape[0])
    return mean, se


class Batch:
    """Systems of uneven size swept together (vmc_batch): one launch per
    kernel per sweep for each group of register-resident systems of one size,
    one persistent launch for the systems on the generic kernel."""

    def __init__(self, systems, walkers, seeds, step0=None, placement=None, stream=None,
                 walker_threads=64):
        n = len(systems)
        self.systems = list(systems)
        self.walkers = [int(w) for w in walkers]
        self.seeds = [int(x) for x in seeds]
        self.stream = stream
        self._descs = (SystemDesc * n)(*[sy.desc() for sy in systems])
        w = (C.c_int64 * n)(*self.walkers)
        sd = (C.c_uint64 * n)(*self.seeds)
        st = np.ascontiguousarray([sy.box_length / 10.0 for sy in systems] if step0 is None else step0,
                                  dtype=np.float64)
        pl = None if placement is None else np.ascontiguousarray(placement, dtype=np.int32)
        self._h = _VP()
        _check(lib().vmc_batch_create(C.cast(self._descs, C.c_void_p), w, sd, _ptr(st),
                                      None if pl is None else _ptr(pl, _ip), n, int(walker_threads),
                                      C.byref(self._h)))

    def close(self):
        if self._h:
            lib().vmc_batch_destroy(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_positions(self):
        _check(lib().vmc_batch_init_positions(self._h, self.stream))

    def sweep(self, n_sweeps=1, measure=False, adapt=False, block_size=1):
        flags = (SWEEP_MEASURE if measure else 0) | (SWEEP_ADAPT if adapt else 0)
        _check(lib().vmc_batch_sweep(self._h, n_sweeps, flags, block_size, self.stream))

    def capture(self, n_sweeps, measure=False, adapt=False, block_size=1):
        flags = (SWEEP_MEASURE if measure else 0) | (SWEEP_ADAPT if adapt else 0)
        _check(lib().vmc_batch_capture(self._h, n_sweeps, flags, block_size))

    def replay(self):
        _check(lib().vmc_batch_replay(self._h, self.stream))

    def member(self, i):
        h = _VP()
        _check(lib().vmc_batch_member(self._h, i, C.byref(h)))
        return Ensemble._borrowed(h, self.systems[i], self.walkers[i], self.seeds[i], self.stream)

    def info(self):
        ns, gr = C.c_int32(), C.c_int32()
        tw, lds = C.c_int64(), C.c_int64()
        _check(lib().vmc_batch_info(self._h, C.byref(ns), C.byref(tw), C.byref(gr), C.byref(lds)))
        return dict(n_systems=ns.value, total_walkers=tw.value, grid=gr.value, lds_bytes=lds.value)

    def set_kernel(self, kernel, system=-1):
        """Kernel of one system (or all): KERNEL_AUTO, KERNEL_GENERIC or
        KERNEL_REGISTER; regroups the batch and drops a captured graph."""
        _check(lib().vmc_batch_set_kernel(self._h, int(system), int(kernel)))

    def set_overlap(self, on):
        _check(lib().vmc_batch_set_overlap(self._h, 1 if on else 0))

    def groups(self):
        """(family of each group, group of each system); family 0 generic,
        1 register, 2 multi-wavefront register."""
        n = len(self.systems)
        ng = C.c_int32()
        fam = np.zeros(n, dtype=np.int32)
        grp = np.zeros(n, dtype=np.int32)
        _check(lib().vmc_batch_groups(self._h, C.byref(ng), _ptr(fam, _ip), _ptr(grp, _ip), n))
        return fam[:ng.value].tolist(), grp.tolist()


def footprint(system: System, n_walkers: int) -> int:
    b = C.c_int64()
    d = system.desc()
    _check(lib().vmc_ensemble_footprint(C.byref(d), int(n_walkers), C.byref(b)))
    return b.value


def autotune(system: System, n_walkers: int, drift_tol=1e-3, sweeps=10, stream=None):
    """Kernel family, variant, inverse placement and (mixed precision) the
    re-inversion interval for this system, chosen by timing on the device."""
    t = Tuning()
    d = system.desc()
    _check(lib().vmc_autotune(C.byref(d), int(n_walkers), float(drift_tol), int(sweeps), C.byref(t),
                              stream))
    return dict(kernel=t.kernel, walker_threads=t.walker_threads, rr_variant=t.rr_variant,
                inv_placement=t.inv_placement,
                reinvert_interval=t.reinvert_interval, moves_per_s=t.moves_per_s, drift=t.drift)


class Server:
    """Continuous-batching request server (vmc_server)."""

    def __init__(self, hbm_budget_bytes, max_active=64, sweeps_per_turn=1):
        self._h = _VP()
        _check(lib().vmc_server_create(int(hbm_budget_bytes), int(max_active), int(sweeps_per_turn),
                                       C.byref(self._h)))

    def submit(self, system: System, n_walkers, warmup_sweeps, measure_sweeps, seed, step0=None,
               block_size=10):
        r = Request(system.desc(), int(n_walkers), int(warmup_sweeps), int(measure_sweeps),
                    int(block_size), int(seed),
                    float(system.box_length / 10.0 if step0 is None else step0))
        i = C.c_int64()
        _check(lib().vmc_server_submit(self._h, C.byref(r), C.byref(i)))
        return i.value

    def poll(self, wait=False, capacity=256):
        buf = (Result * capacity)()
        n = C.c_int32()
        _check(lib().vmc_server_poll(self._h, buf, capacity, C.byref(n), 1 if wait else 0))
        return [buf[i].as_dict() for i in range(n.value)]

    def cancel(self, request_id):
        """Ends a request: its result reports status ERR_CANCELLED."""
        _check(lib().vmc_server_cancel(self._h, int(request_id)))

    def close(self):
        if self._h:
            _check(lib().vmc_server_destroy(self._h))
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_walker_threads(system: System) -> int:
    d = system.desc()
    return lib().vmc_default_walker_threads(C.byref(d))
