"""Pair-kernel shapes (the autotuner's choices: threads per block, and
per-neighbour data gathered or staged in LDS) change no result, and the
autotuner leaves the state as it found it."""
import ctypes as C

import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu


def _run(parallel, case, shape, steps):
    from eppx import Algorithm, SPH
    ps, ns = U.device(parallel, case, 64)
    Algorithm.set_pair_shape(ps, *shape)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    Algorithm.step(ps, ns, case.physics, n_steps=steps)
    host = U.download(ps)
    ps.close()
    ns.close()
    return host


def test_shapes_give_identical_results(parallel):
    """FAST math, 8 steps of the dx = 0.01 dam break with every shape."""
    from eppx import Algorithm, cases
    case = cases.dam_break(0.01, 2.0, M=64)
    parallel.set_math("fast")
    try:
        shapes = [((256, 0), (256, 0))] + [(a, b) for a, b in zip(Algorithm.SHAPES[1:], Algorithm.SHAPES[::-1])]
        runs = [_run(parallel, case, sh, 8) for sh in shapes]
    finally:
        parallel.set_math("exact")
    for other in runs[1:]:
        assert np.array_equal(runs[0].is_alive_, other.is_alive_)
        assert np.array_equal(runs[0].float_properties_, other.float_properties_)
        assert np.array_equal(runs[0].int_properties_, other.int_properties_)


def test_autotune_leaves_state_unchanged(parallel):
    from eppx import Algorithm, SPH, cases
    case = cases.dam_break(0.01, 2.0, M=64)
    parallel.set_math("fast")
    try:
        ps, ns = U.device(parallel, case, 64)
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(case.physics))
        Algorithm.step(ps, ns, case.physics, n_steps=3)
        before = U.download(ps)
        total0 = C.c_int64()
        from eppx import _lib
        _lib.check(_lib.lib().eppx_pair_total(ps.handle, C.byref(total0), 0), parallel.ctx, "total")
        choice = Algorithm.autotune(ps, ns, case.physics, trials=2)
        assert choice["search"]["block"] in (256, 512) and choice["momentum"]["block"] in (256, 512)
        assert all(v > 0 for v in choice["search_us"].values())
        assert all(v > 0 for v in choice["momentum_us"].values())
        after = U.download(ps)
        total1 = C.c_int64()
        _lib.check(_lib.lib().eppx_pair_total(ps.handle, C.byref(total1), 0), parallel.ctx, "total")
        assert total0.value == total1.value
        assert np.array_equal(before.float_properties_, after.float_properties_)
        assert np.array_equal(before.int_properties_, after.int_properties_)
        ps.close()
        ns.close()
    finally:
        parallel.set_math("exact")
