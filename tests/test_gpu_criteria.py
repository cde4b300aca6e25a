"""Pair sets on the search paths the plain dam break does not reach, against
the restatement, in EXACT and FAST mode (the criterion is exact in both):

* particles with different H: the symmetry criterion's harmonic-mean test
  decides every candidate between the bounds set from the H range;
* a third tag with partial adjacency (3-3 and 3-2 not adjacent): the
  per-candidate tag test, and counts that include pairs the continuity
  action does not visit;
* defaultSearchCriterion (dr2 <= gap^2).

Checked per particle: nCount of the search kernel (the CSR offsets), the
neighbour list in list order, and the fused continuity sum dDensity.
"""
import numpy as np
import pytest

import oracle as O
import parity_util as U

pytestmark = pytest.mark.gpu

TAG, POS, H = 0, 0, 14  # column indices of the dam-break layout (Tag; PositionVec; H)


def _case(seed, vary_h=False, third_tag=False):
    from eppx import cases
    case = cases.dam_break(0.01, 2.0, M=64)
    rng = np.random.default_rng(seed)
    h = case.host
    fluid = h.int_properties_[:, TAG] == 1
    pos = h.float_properties_[:, POS:POS + 2]
    pos[fluid] += (rng.uniform(-0.45, 0.45, (fluid.sum(), 2)) * case.dx).astype(np.float32)
    if vary_h:  # H in [0.8, 1] h0: every pair still fits the cells of size 2 h0
        h.float_properties_[:, H] *= rng.uniform(0.8, 1.0, h.n_capacity).astype(np.float32)
    if third_tag:
        pick = fluid & (rng.uniform(size=h.n_capacity) < 0.3)
        h.int_properties_[pick, TAG] = 3
    return case


def _both(par, case, pairs, criterion):
    from eppx import Algorithm, Class, SPH, cases
    ps = Class.ParticleSystem(par, case.domain, case.n_particles, case.int_nt, case.float_nt,
                              case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(par, case.domain, pairs, max_neighbour_number=64)
    crit = Algorithm.symmetrySearchCriterion if criterion == O.CRIT_SYMMETRY else Algorithm.defaultSearchCriterion
    Algorithm.search(ps, case.domain, ns, criterion=crit, action=SPH.sComputeKernelAndiContinuity(case.physics))
    h = case.host
    adj = Class.ActivePair(pairs).adjacency_matrix_
    ref = O.System(h.int_properties_.copy(order="F"), h.float_properties_.copy(order="F"), h.is_alive_.copy(),
                   h.cell_index_.copy(), h.n_particles_, O.make_layout(case.int_nt, case.float_nt, h.n_capacity),
                   O.domain2d(*case.domain.args), adj, 64)
    assert ref.search(criterion, 1, U.oracle_physics(case.physics)) == 0
    return ps, ns, ref


def _check(ps, ns, ref):
    off, js = U.export_pairs(ps, ns)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert off[-1] == roff[-1]
    assert np.array_equal(off, roff), "per-particle neighbour counts differ"
    assert np.array_equal(js, rjs), "neighbour lists differ"
    return off


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_varying_h_pairs(parallel, math):
    from eppx import cases
    case = _case(11, vary_h=True)
    parallel.set_math(math)
    try:
        ps, ns, ref = _both(parallel, case, cases.ACTIVE_PAIR, O.CRIT_SYMMETRY)
        _check(ps, ns, ref)
        if math == "exact":
            U.assert_state_equal(U.download(ps), ref, names=("dDensity",))
    finally:
        parallel.set_math("exact")


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_third_tag_partial_adjacency(parallel, math):
    pairs = [(1, 1), (1, 2), (2, 1), (1, 3), (3, 1)]
    case = _case(12, third_tag=True)
    parallel.set_math(math)
    try:
        ps, ns, ref = _both(parallel, case, pairs, O.CRIT_SYMMETRY)
        off = _check(ps, ns, ref)
        tags = case.host.int_properties_[: case.n_particles, TAG]
        assert (np.diff(off)[tags == 3] > 0).all()  # tag-3 particles see fluid neighbours
        if math == "exact":
            U.assert_state_equal(U.download(ps), ref, names=("dDensity",))
    finally:
        parallel.set_math("exact")


def test_default_criterion_pairs(parallel):
    from eppx import cases
    case = _case(13)
    ps, ns, ref = _both(parallel, case, cases.ACTIVE_PAIR, O.CRIT_DEFAULT)
    _check(ps, ns, ref)
    U.assert_state_equal(U.download(ps), ref, names=("dDensity",))
