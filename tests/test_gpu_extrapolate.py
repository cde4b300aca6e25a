"""Extrapolated wall pressure (iExtrapolatePressure! / sExtrapolatePressure!,
Pressure.jl:81-143) on the device against the restatement, bit for bit, and
the filter / extrapolation sums carried through steps."""
import numpy as np
import pytest

import oracle as O
import parity_util as U

pytestmark = pytest.mark.gpu

SUMS = ("SumWeight", "SumWeightedDensity", "SumWeightedPressure")


@pytest.fixture(scope="module")
def small():
    from eppx import cases
    return cases.small(M=40)


def test_extrapolated_wall_pressure_exact(parallel, small):
    """25 steps, then the extrapolation after a search, then 5 more steps (the
    permute now carries the sums) -- equal to the restatement throughout."""
    from eppx import Algorithm, SPH
    ph = small.physics
    ps, ns, ref = U.search_both(parallel, small)
    Algorithm.step(ps, ns, ph, n_steps=25)
    err, _ = ref.run_steps(25, U.oracle_physics(ph))
    assert err == 0
    Algorithm.interaction(ps, SPH.iExtrapolatePressure(ph))
    Algorithm.selfaction(ps, SPH.sExtrapolatePressure(ph))
    ref.interaction(3, U.oracle_physics(ph))
    ref.selfaction(4, U.oracle_physics(ph))
    host = U.download(ps)
    U.assert_state_equal(host, ref, names=U.STATE_FLOAT + SUMS)
    wall = host.int_properties_[: ref.n, ref.L.Tag] == 2
    assert (host.float_properties_[: ref.n, ref.L.Pressure][wall] > 0.0).sum() > 100  # walls under the column
    # interaction only: the sums stay in the state and move with their particles
    Algorithm.interaction(ps, SPH.iExtrapolatePressure(ph))
    ref.interaction(3, U.oracle_physics(ph))
    Algorithm.step(ps, ns, ph, n_steps=5)
    err, _ = ref.run_steps(5, U.oracle_physics(ph))
    assert err == 0
    U.assert_state_equal(U.download(ps), ref, names=U.STATE_FLOAT + SUMS)


def test_sums_uploaded_non_zero_survive_steps(parallel, small):
    """Sums that arrive non-zero are part of the state: steps without a filter
    leave them as they were, row by row."""
    from eppx import Algorithm, Class, SPH, cases
    rng = np.random.default_rng(5)
    case = cases.small(M=40)
    h = case.host
    for name in SUMS:
        c = case.host.named_index_.float_table.index[name] - 1
        h.float_properties_[: h.n_particles_, c] = rng.uniform(0.1, 2.0, h.n_particles_).astype(np.float32)
    ps, ns, ref = U.search_both(parallel, case)
    Algorithm.step(ps, ns, case.physics, n_steps=10)
    err, _ = ref.run_steps(10, U.oracle_physics(case.physics))
    assert err == 0
    U.assert_state_equal(U.download(ps), ref, names=U.STATE_FLOAT + SUMS)
