"""A neighbour system built with a periodic boundary policy, on the device.

The reference allocates the periodic shift table and leaves filling it as a
TODO (PeriodicBoundary.jl:57), so the shift applyPeriodicBoundaryPolicy adds
to (dx, dy) (NeighbourSearch.jl:129-139) is zero and a periodic policy gives
the non-periodic pairs and sums.  Here: a 2D dam break built with a periodic
policy runs bit-identically to the restatement (EXACT), 3D likewise against a
run with no periodic policy, and a configuration written from the device
neighbour system carries the policy back.
"""
import numpy as np
import pytest

import oracle as O
import parity_util as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", ["PeriodicBoundaryPolicy2DAlongX", "PeriodicBoundaryPolicy2DAlongXY"])
def test_periodic_policy_2d_steps_match_restatement(parallel, policy):
    from eppx import Algorithm, Class, SPH, cases
    case = cases.small()
    ps = Class.ParticleSystem(parallel, case.domain, case.n_particles, case.int_nt, case.float_nt,
                              case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(parallel, case.domain, cases.ACTIVE_PAIR, getattr(Class, policy),
                               max_neighbour_number=case.M)
    assert ns.periodic_boundary_.neighbour_cell_relative_position_list_.shape == (case.domain.get_n(), 9, 2)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    n = 25
    Algorithm.step(ps, ns, case.physics, n_steps=n)
    host = U.download(ps)
    ref = U.reference(case)
    oph = U.oracle_physics(case.physics)
    assert ref.search(O.CRIT_SYMMETRY, 1, oph) == 0
    err, _ = ref.run_steps(n, oph)
    assert err == 0
    U.assert_state_equal(host, ref)
    assert np.array_equal(host.cell_index_[: ref.n], ref.cell_index[: ref.n])
    ps.close()
    ns.close()


def test_periodic_policy_3d_matches_none(parallel):
    from eppx import Algorithm, Class, SPH, cases
    case = cases.dam_break_3d(0.05, M=300, block=(0.5, 0.25, 0.5), tank=(1.2, 0.5, 0.8))
    out = []
    for policy in (Class.NonePeriodicBoundaryPolicy, Class.PeriodicBoundaryPolicy3DAlongXYZ):
        ps = Class.ParticleSystem(parallel, case.domain, case.n_particles, case.int_nt, case.float_nt,
                                  case.basic_parameters)
        Class.syncto(ps, case.host)
        ns = Class.NeighbourSystem(parallel, case.domain, cases.ACTIVE_PAIR, policy, max_neighbour_number=300)
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(case.physics))
        Algorithm.step(ps, ns, case.physics, n_steps=10)
        out.append(U.download(ps))
        ps.close()
        ns.close()
    a, b = out
    assert np.array_equal(a.is_alive_, b.is_alive_)
    assert np.array_equal(a.int_properties_, b.int_properties_)
    assert np.array_equal(a.float_properties_.view(np.uint32), b.float_properties_.view(np.uint32))


def test_periodic_policy_through_configuration(parallel, tmp_path):
    from eppx import Class, DataIO, cases
    case = cases.small()
    ps, _ = U.device(parallel, case)
    ns = Class.NeighbourSystem(parallel, case.domain, cases.ACTIVE_PAIR, Class.PeriodicBoundaryPolicy2DAlongY,
                               max_neighbour_number=case.M)
    w = DataIO.Writer(str(tmp_path))
    c2 = DataIO.parse(DataIO.writeConfig(w, DataIO.config(parallel, case.domain, ps, ns, w)))
    assert c2["neighbour_system"]["periodic_boundary_policy"]["type"] == "2D"
    ns2 = DataIO.NeighbourSystem(c2, parallel, DataIO.Domain(c2))
    assert ns2.periodic_boundary_.policy is Class.PeriodicBoundaryPolicy2DAlongY
    for h in (ps, ns, ns2):
        h.close()
