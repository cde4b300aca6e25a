"""Kernel paths the dam-break cases never take, against the restatement.

* Cells much larger than 2h (domain gap 5 dx, 25 particles per cell): every
  neighbour row spans more than 64 slots, so every particle walk tests its
  candidates one at a time from HBM instead of through the accept masks
  (pair_device.h, walk_rows).  EXACT steps bit-identical; FAST within the
  per-particle bars of the other FAST tests; the walks counted by
  eppx_pair_row_fallback.
* A cell with more than 1024 particles: the in-cell order pass gives up, and
  the search reports EPPX_E_CAPACITY ("a cell holds more than 1024
  particles"); the pair sets stay exact.
"""
import ctypes as C

import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu


def row_fallback(ps, reset=True):
    from eppx import _lib
    v = C.c_int64()
    _lib.check(_lib.lib().eppx_pair_row_fallback(ps.handle, C.byref(v), int(reset)), ps.parallel.ctx, "row fallback")
    return v.value


def wide_cells_case():
    from eppx import Class, cases
    case = cases.small(M=40)
    g = 5 * case.dx  # cells of 5 dx: 3 cells of a neighbour row hold 75 lattice slots
    case.domain = Class.Domain2D(g, -g, -g, 4.0 + g, 3.0 + g)
    return case


def test_wide_cells_exact(parallel):
    from eppx import Algorithm
    case = wide_cells_case()
    ph = case.physics
    ps, ns, ref = U.search_both(parallel, case)
    row_fallback(ps)  # the first search's walks
    Algorithm.step(ps, ns, ph, n_steps=5)
    walks = row_fallback(ps)
    err, pairs = ref.run_steps(5, U.oracle_physics(ph))
    assert err == 0
    host = U.download(ps)
    U.assert_state_equal(host, ref, alive_only=False)
    U.assert_neighbour_columns_equal(host, ref)
    off, js = U.export_pairs(ps, ns)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    n_fluid = int(((ref.is_alive[: ref.n] == 1) & (ref.IP[: ref.n, ref.L.Tag] == 1)).sum())
    print(f"wide cells: {walks} walks through HBM over 5 steps ({n_fluid} fluid particles)")
    # each step walks the fluid twice (search, momentum); all but the column's edge cells take this path
    assert walks >= 5 * n_fluid


def test_wide_cells_fast(parallel):
    from eppx import Algorithm
    case = wide_cells_case()
    ph = case.physics
    parallel.set_math("fast")
    try:
        ps, ns, ref = U.search_both(parallel, case)
        Algorithm.step(ps, ns, ph, n_steps=5)
        walks = row_fallback(ps)
        host = U.download(ps)
        off, js = U.export_pairs(ps, ns)
    finally:
        parallel.set_math("exact")
    err, _ = ref.run_steps(5, U.oracle_physics(ph))
    assert err == 0
    assert walks > 0
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    L = ref.L
    alive = ref.is_alive[: ref.n] == 1
    g, r = host.float_properties_[: ref.n][alive], ref.FP[: ref.n][alive]
    rho = float(np.max(np.abs(g[:, L.Density].astype(np.float64) - r[:, L.Density]) / np.abs(r[:, L.Density])))
    t = 5 * float(np.float32(ph.dt))
    vr = r[:, L.VelocityVec: L.VelocityVec + 2].astype(np.float64)
    vel = float(np.max(np.abs(g[:, L.VelocityVec: L.VelocityVec + 2] - vr) / (1e-5 * np.abs(vr) + 1e-5 * 9.81 * t)))
    print(f"wide cells FAST: density {rho:.3g}, velocity {vel:.3g} of the bar")
    assert rho <= 1e-5
    assert vel <= 1.0


def crowded_cell_case(n_side=34, M=2000):
    """n_side^2 fluid particles on a 1 mm lattice moved into the middle 40 mm
    cell of a 3 x 3 grid (the fluid rows of a small dam break, which come
    first; its wall rows are dropped), H = 20 mm."""
    from eppx import Class, SPH, cases
    case = cases.dam_break(0.001, 2.0, SPH.WendlandC2, M=M, fluid_width=0.001 * n_side,
                           fluid_height=0.001 * n_side, box_width=0.04, box_height=0.04)
    h = case.host
    n = n_side * n_side
    assert (h.int_properties_[:n, h.named_index_.int_table.index["Tag"] - 1] == 1).all()
    # keep the fluid rows only
    h.n_capacity = h.n_particles_ = n
    h.int_properties_ = np.asfortranarray(h.int_properties_[:n])
    h.float_properties_ = np.asfortranarray(h.float_properties_[:n])
    h.is_alive_ = h.is_alive_[:n].copy()
    h.cell_index_ = h.cell_index_[:n].copy()
    x = h.named_index_.float_table.index["PositionVec"] - 1
    h.float_properties_[:n, x] += np.float32(0.04)
    h.float_properties_[:n, x + 1] += np.float32(0.04)
    h.float_properties_[:n, h.named_index_.float_table.index["H"] - 1] = np.float32(0.02)
    case.domain = Class.Domain2D(0.04, 0.0, 0.0, 0.12, 0.12)
    return case


def test_crowded_cell_reports_capacity(parallel):
    from eppx import Algorithm, EppxError, SPH
    case = crowded_cell_case()
    assert case.n_particles > 1024
    ps, ns = U.device(parallel, case)
    with pytest.raises(EppxError) as e:
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(case.physics))
    assert e.value.status == 3 and "more than 1024" in str(e.value)
    parallel.synchronize()  # reported once, then cleared
    # the pair sets are still the reference's (only the in-cell order is unspecified)
    ref = U.reference(case)
    import oracle as O
    assert ref.search(O.CRIT_SYMMETRY, 1, U.oracle_physics(case.physics)) == 0
    off, js = U.export_pairs(ps, ns)
    roff, rjs = ref.pair_csr()
    assert np.array_equal(off, roff)
    got = np.concatenate([np.sort(js[off[i]: off[i + 1]]) for i in range(case.n_particles)])
    assert np.array_equal(got, rjs)
    assert row_fallback(ps) > 0
