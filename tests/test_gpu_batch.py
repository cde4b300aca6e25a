"""The batch (many cases in one segmented cell list, one launch per kernel
per step) against the restatement run case by case, and against the plain
particle-system path: each case's result must be the same alone and in the
mix.  EXACT math: bit-identical to the restatement.  FAST math: bit-identical
to the same case run alone on the plain GPU path (the per-case summation
order does not depend on the other cases)."""
import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu

SPECS = [(0.02, 1.0, 2.0, 12), (0.02, 0.7, 1.1, 7), (0.04, 1.3, 0.6, 20)]  # dx, width, height, steps


def _case(dx, w, h, steps):
    from eppx import SPH, cases
    return cases.dam_break(dx, 2.0, SPH.WendlandC2, M=64, fluid_width=w, fluid_height=h, steps=steps)


def _batch(parallel, cs, capacity=None, case_capacity=8):
    from eppx import Batch, cases
    cap = capacity if capacity is not None else sum(c.n_particles for c in cs) + 1000
    return Batch.Batch(parallel, cs[0].host, cap, case_capacity, cases.ACTIVE_PAIR, max_neighbour_number=64)


def _oracle_after(case, steps):
    ref = U.reference(case, 64)
    oph = U.oracle_physics(case.physics)
    assert ref.search(U.O.CRIT_SYMMETRY, 1, oph) == 0
    err, _ = ref.run_steps(steps, oph)
    assert err == 0
    return ref


def _assert_case(host, ref):
    U.assert_state_equal(host, ref)
    assert np.array_equal(host.cell_index_[: ref.n], ref.cell_index[: ref.n])


def test_batch_cases_match_restatement(parallel):
    """Three uneven cases with budgets 12, 7 and 20 in one batch; 20 batch steps."""
    cs = [_case(*s) for s in SPECS]
    b = _batch(parallel, cs)
    try:
        ids = [b.add(c.domain, c.physics, c.host, c.steps) for c in cs]
        b.step(20)
        assert (b.steps_left()[ids] == 0).all()
        for cid, c in zip(ids, cs):
            _assert_case(b.download(cid), _oracle_after(c, c.steps))
    finally:
        b.close()


def test_batch_add_and_remove_between_steps(parallel):
    """Cases join a running batch, finish at their own budgets, stay frozen
    while others run, and leave; the freed slots take a new case."""
    a, bb, c = _case(0.02, 1.0, 2.0, 10), _case(0.02, 0.6, 0.9, 6), _case(0.04, 1.2, 1.0, 5)
    b = _batch(parallel, [a, bb, c], capacity=a.n_particles + bb.n_particles + 200)
    try:
        ia = b.add(a.domain, a.physics, a.host, 10)
        b.step(4)
        ib = b.add(bb.domain, bb.physics, bb.host, 6)
        b.step(10)
        left = b.steps_left()
        assert left[ia] == 0 and left[ib] == 0
        _assert_case(b.download(ia), _oracle_after(a, 10))
        _assert_case(b.download(ib), _oracle_after(bb, 6))
        b.remove(ia)
        assert b.info() == (bb.n_particles, 1)
        ic = b.add(c.domain, c.physics, c.host, 5)  # fits only in the slots case a released
        b.step(7)
        _assert_case(b.download(ib), _oracle_after(bb, 6))
        _assert_case(b.download(ic), _oracle_after(c, 5))
    finally:
        b.close()


def test_batch_remove_middle_case_then_step(parallel):
    """Removing a case that is not the last moves the later cases to lower
    cells: the steps after it (with no add between) walk the cell list of the
    new layout and the remaining cases go on as if alone."""
    cs = [_case(*s) for s in SPECS]  # budgets 12, 7, 20
    b = _batch(parallel, cs)
    try:
        ids = [b.add(c.domain, c.physics, c.host, 20) for c in cs]
        b.step(5)
        b.remove(ids[1])
        assert b.info() == (cs[0].n_particles + cs[2].n_particles, 2)
        b.step(6)
        left = b.steps_left()
        assert left[ids[0]] == 9 and left[ids[2]] == 9
        _assert_case(b.download(ids[0]), _oracle_after(cs[0], 11))
        _assert_case(b.download(ids[2]), _oracle_after(cs[2], 11))
        b.remove(ids[0])
        b.step(3)
        _assert_case(b.download(ids[2]), _oracle_after(cs[2], 14))
    finally:
        b.close()


def test_batch_fast_equals_alone(parallel):
    """FAST math: every case of the batch equals the same case run alone
    through the particle-system path (tolerance: none; identical bits)."""
    from eppx import Algorithm, Class, SPH, cases
    cs = [_case(*s) for s in SPECS]
    parallel.set_math("fast")
    try:
        b = _batch(parallel, cs)
        try:
            ids = [b.add(c.domain, c.physics, c.host, c.steps) for c in cs]
            b.step(20)
            mixed = [b.download(i) for i in ids]
        finally:
            b.close()
        for c, got in zip(cs, mixed):
            ps = Class.ParticleSystem(parallel, c.domain, c.n_particles, c.int_nt, c.float_nt, c.basic_parameters)
            Class.syncto(ps, c.host)
            ns = Class.NeighbourSystem(parallel, c.domain, cases.ACTIVE_PAIR, max_neighbour_number=64)
            Algorithm.search(ps, c.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=SPH.sComputeKernelAndiContinuity(c.physics))
            Algorithm.step(ps, ns, c.physics, n_steps=c.steps)
            alone = U.download(ps)
            n = c.n_particles
            assert np.array_equal(got.is_alive_[:n], alone.is_alive_[:n])
            L = ps.layout
            for col in (L.PositionVec, L.PositionVec + 1, L.VelocityVec, L.VelocityVec + 1, L.Density, L.Pressure,
                        L.dDensity, L.Volume):
                assert np.array_equal(got.float_properties_[:n, col], alone.float_properties_[:n, col])
            ps.close()
            ns.close()
    finally:
        parallel.set_math("exact")


def test_batch_rejects_what_does_not_fit(parallel):
    from eppx import EppxError, SPH, cases
    a = _case(0.02, 1.0, 2.0, 3)
    b = _batch(parallel, [a], capacity=a.n_particles - 1)
    try:
        with pytest.raises(EppxError) as e:
            b.add(a.domain, a.physics, a.host, 3)
        assert e.value.status == 3  # EPPX_E_CAPACITY
        other = cases.dam_break(0.02, 1.5, SPH.CubicSpline, M=64)  # another kernel: not the batch's physics
        small = _batch(parallel, [a])
        try:
            small.add(a.domain, a.physics, a.host, 1)
            with pytest.raises(EppxError) as e2:
                small.add(other.domain, other.physics, other.host, 1)
            assert e2.value.status == 1  # EPPX_E_ARG
        finally:
            small.close()
    finally:
        b.close()
