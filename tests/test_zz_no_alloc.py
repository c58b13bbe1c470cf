"""vmc_ensemble_measure_noalloc against vmc_ensemble_measure.

The allocation-free entry point stages the per-walker parts in the library's
per-device array instead of a fresh allocation; it runs the same kernels, so
every part must come out bit for bit the same as vmc_ensemble_measure's, on
consecutive calls whose inputs shrink (the array keeps the previous call's
contents beyond the new walker count), on every kernel family and precision.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # (N, r_s, b, precision, K, walkers, kernel) in call order
    (54, 2.0, 1.0, 1, 256, 512, None),   # register, mixed
    (54, 2.0, 1.0, 1, 256, 40, None),    # the same size with fewer walkers
    (14, 1.0, 1.0, 0, 0, 9, None),       # register, FP64, smaller system
    (66, 3.0, 1.2, 1, 64, 33, None),     # multi-wavefront register, mixed
    (66, 3.0, 1.2, 0, 0, 12, None),      # multi-wavefront register, FP64
    (38, 2.0, 0.8, 1, 32, 21, 1),        # generic kernel, mixed
    (38, 2.0, 0.8, 0, 0, 7, 1),          # generic kernel, FP64
]


def run_case(vmc, N, rs, b, prec, K, W, kernel, seed):
    s = vmc.System.heg(N, rs, b, precision=prec, reinvert_interval=K)
    e = vmc.Ensemble(s, W, seed)
    try:
        if kernel is not None:
            e.set_kernel(kernel)
        e.init_positions()
        e.sweep(2, adapt=True)
        e.sweep(1, measure=True)
        ref = e.measure()
        got1 = e.measure_noalloc()
        got2 = e.measure_noalloc()
    finally:
        e.close()
    return ref, got1, got2


def test_noalloc_measure_equals_measure(vmc):
    for i, (N, rs, b, prec, K, W, kernel) in enumerate(CASES):
        print(f"case {i}: N={N} precision={prec} K={K} W={W} kernel={kernel}", flush=True)
        ref, got1, got2 = run_case(vmc, N, rs, b, prec, K, W, kernel, 9000 + i)
        for k in vmc.PART:
            assert ref[k].shape == (W,)
            np.testing.assert_array_equal(got1[k], ref[k], err_msg=f"case {i} part {k}")
            np.testing.assert_array_equal(got2[k], ref[k], err_msg=f"case {i} part {k} (repeat)")
        assert np.all(np.isfinite(ref["ELOCAL"]))


def test_noalloc_measure_rejects_more_walkers_than_its_array(vmc):
    s = vmc.System.heg(14, 1.0, 1.0)
    e = vmc.Ensemble(s, vmc.NOALLOC_MAX_WALKERS + 1, 5)
    try:
        with pytest.raises(vmc.VmcError) as err:
            e.measure_noalloc()
        assert err.value.code == vmc.ERR_INVALID
        assert "at most 65536 walkers" in str(err.value)
    finally:
        e.close()
