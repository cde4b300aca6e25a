"""Helpers shared by the GPU parity tests: the same case on the device (through
the C ABI) and in the restatement (the checker)."""
import ctypes as C

import numpy as np

import oracle as O
from eppx import Algorithm, Class, SPH, _lib, cases

ADJ = np.array([[1, 1], [1, 0]], np.int32)
STATE_FLOAT = ("PositionVec", "VelocityVec", "dVelocityVec", "Mass", "Volume", "Density", "dDensity", "Pressure",
               "H")


def device(par, case, M=None):
    M = M or case.M
    ps = Class.ParticleSystem(par, case.domain, case.n_particles, case.int_nt, case.float_nt,
                              case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(par, case.domain, cases.ACTIVE_PAIR, max_neighbour_number=M)
    return ps, ns


def reference(case, M=None):
    M = M or case.M
    h = case.host
    L = O.make_layout(case.int_nt, case.float_nt, h.n_capacity)
    return O.System(h.int_properties_.copy(order="F"), h.float_properties_.copy(order="F"), h.is_alive_.copy(),
                    h.cell_index_.copy(), h.n_particles_, L, O.domain2d(*case.domain.args), ADJ, M)


def oracle_physics(physics):
    return O.physics(**physics.as_dict())


def search_both(par, case, M=None):
    ps, ns = device(par, case, M)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    ref = reference(case, M)
    assert ref.search(O.CRIT_SYMMETRY, 1, oracle_physics(case.physics)) == 0
    return ps, ns, ref


def download(ps):
    host = Class.mirror(ps)
    Class.syncto(host, ps)
    return host


def width(name):
    return 2 if name.endswith("Vec") else 1


def assert_state_equal(host, ref, names=STATE_FLOAT, alive_only=True):
    assert np.array_equal(host.is_alive_[: ref.n], ref.is_alive[: ref.n])
    rows = ref.is_alive[: ref.n] == 1 if alive_only else slice(0, ref.n)
    for name in names:
        c = getattr(ref.L, name)
        got = host.float_properties_[: ref.n][rows, c: c + width(name)]
        exp = ref.FP[: ref.n][rows, c: c + width(name)]
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            r, k = bad[0]
            raise AssertionError(f"{name}: {len(bad)} entries differ, first row {r}: {got[r, k]!r} vs {exp[r, k]!r}")


def assert_neighbour_columns_equal(host, ref):
    """nCount, nIndex in list order, nRVec, nR, nW, nDW, nHInv (alive rows)."""
    L = ref.L
    M = L.max_neighbour_number
    alive = np.nonzero(ref.is_alive[: ref.n] == 1)[0]
    cnt = ref.IP[alive, L.nCount]
    assert np.array_equal(host.int_properties_[alive, L.nCount], cnt)
    mask = np.arange(M)[None, :] < cnt[:, None]
    for col, w, table in (("nIndex", 1, "int"), ("nR", 1, "f"), ("nW", 1, "f"), ("nDW", 1, "f"),
                          ("nHInv", 1, "f"), ("nRVec", 2, "f")):
        c = getattr(L, col)
        got = (host.int_properties_ if table == "int" else host.float_properties_)[alive, c: c + w * M]
        exp = (ref.IP if table == "int" else ref.FP)[alive, c: c + w * M]
        m = np.repeat(mask, w, axis=1)
        if not np.array_equal(got[m], exp[m]):
            bad = np.argwhere((got != exp) & m)
            raise AssertionError(f"{col}: {len(bad)} entries differ (first {bad[0]})")


def export_pairs(ps, ns):
    L = _lib.lib()
    n = ps.n_particles_
    off = np.zeros(n + 1, np.int64)
    total = C.c_int64()
    _lib.check(L.eppx_pair_count(ps.handle, ns.handle, C.byref(total)), ps.parallel.ctx, "pair_count")
    js = np.zeros(max(1, total.value), np.int32)
    _lib.check(L.eppx_export_pairs(ps.handle, ns.handle, _lib.ptr(off, C.c_int64), _lib.ptr(js, C.c_int32),
                                   js.shape[0]), ps.parallel.ctx, "export_pairs")
    return off, js[: total.value]


def reference_pairs_in_order(ref):
    """CSR of the restatement's neighbour lists in list order (0-based)."""
    L = ref.L
    cnt = np.where(ref.is_alive[: ref.n] == 1, ref.IP[: ref.n, L.nCount], 0).astype(np.int64)
    off = np.zeros(ref.n + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    M = L.max_neighbour_number
    lists = ref.IP[: ref.n, L.nIndex: L.nIndex + M] - 1
    mask = np.arange(M)[None, :] < cnt[:, None]
    return off, lists[mask].astype(np.int32)


def per_particle_bar(got, exp, floor, rel=1e-5):
    """FAST-mode bar per entry: |got - exp| <= rel |exp| + floor.  Returns the
    largest |got - exp| / (rel |exp| + floor) (<= 1 passes) and how many
    entries have |exp| below floor / rel, where the absolute floor dominates."""
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    return (float(np.max(np.abs(got - exp) / (rel * np.abs(exp) + floor), initial=0.0)),
            int(np.sum(np.abs(exp) < floor / rel)))


def developed(parallel, case, n_steps, math="exact", filter_interval=0):
    """The case after n_steps of its loop on the device: a flow well past its
    start (pressure built up, the column collapsing), used as the common input
    of the FAST-against-restatement comparisons.  Returns a copy of the case
    whose host tables hold that state, with dDensity zeroed: the search that
    both sides run first recomputes the pending dDensity from the same
    positions and velocities."""
    import dataclasses
    parallel.set_math(math)
    try:
        ps, ns = device(parallel, case)
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(case.physics))
        Algorithm.step(ps, ns, case.physics, n_steps=n_steps, filter_interval=filter_interval)
        host = download(ps)
        ps.close()
        ns.close()
    finally:
        parallel.set_math("exact")
    host.float_properties_[:, host.named_index_.float_table.index["dDensity"] - 1] = 0.0
    return dataclasses.replace(case, host=host)


def bar_ratio(got, exp, bound):
    """max |got - exp| / bound over the entries (<= 1 passes)."""
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    d = np.abs(got - exp)
    bound = np.asarray(bound, np.float64)
    ok = bound > 0
    if (d[~ok] > 0).any():
        return float("inf")
    return float(np.max(d[ok] / bound[ok], initial=0.0))


def fast_momentum_pass(parallel, case, shapes=None):
    """FAST math against the restatement from the same tables, per particle.

    (a) A FAST search (kernel + continuity) and sContinuity!: dDensity and
        density.
    (b) An EXACT search and sContinuity! (bit-identical to the restatement's),
        then iMomentum! in FAST math: the acceleration from identical inputs.
    A float32 sum evaluated in another order or with other roundings of its
    terms (the device's row order, FMA, hardware reciprocals) differs from the
    restatement's by a small multiple of eps times the magnitude bound of the
    sum (each term with its inner sums of products in absolute value; oracle
    or_pair_term_magnitude).  The bars are 1e-5 of that bound for dDensity
    and acceleration, 1e-5 relative for density.  For scale, the restatement's
    momentum pass is also summed in reversed neighbour order (exact
    arithmetic, only the order differs -- the reference's GPU lists have no
    fixed order) and measured against the same bar.  Returns the
    measurements."""
    ph = case.physics
    oph = oracle_physics(ph)
    try:
        parallel.set_math("fast")
        ps, ns = device(parallel, case)
        if shapes is not None:
            Algorithm.set_pair_shape(ps, search=shapes["search"], momentum=shapes["momentum"])
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(ph))
        after_search = download(ps)
        Algorithm.selfaction(ps, SPH.sContinuity(ph))
        after_continuity = download(ps)
        ps.close()
        ns.close()
        parallel.set_math("exact")
        ps, ns = device(parallel, case)
        if shapes is not None:
            Algorithm.set_pair_shape(ps, search=shapes["search"], momentum=shapes["momentum"])
        Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(ph))
        Algorithm.selfaction(ps, SPH.sContinuity(ph))
        parallel.set_math("fast")
        Algorithm.interaction(ps, SPH.iMomentum(ph))
        after_momentum = download(ps)
        ps.close()
        ns.close()
    finally:
        parallel.set_math("exact")
    ref = reference(case)
    assert ref.search(O.CRIT_SYMMETRY, 1, oph) == 0
    L = ref.L
    n = ref.n
    alive = ref.is_alive[:n] == 1
    fluid = alive & (ref.IP[:n, L.Tag] == ph.fluid_tag)
    drho = bar_ratio(after_search.float_properties_[:n][alive, L.dDensity], ref.FP[:n][alive, L.dDensity],
                     1e-5 * ref.term_magnitude(0, oph)[alive])
    ref.selfaction(1, oph)
    rho_g = after_continuity.float_properties_[:n][alive, L.Density].astype(np.float64)
    rho_r = ref.FP[:n][alive, L.Density].astype(np.float64)
    a_m = ref.term_magnitude(1, oph)[fluid]
    du_c = slice(L.dVelocityVec, L.dVelocityVec + 2)
    before = ref.FP[:, du_c].copy()
    ref.set_pair_order(1)
    ref.interaction(1, oph)
    du_rev = ref.FP[:n][fluid, du_c].copy()
    ref.FP[:, du_c] = before
    ref.set_pair_order(0)
    ref.interaction(1, oph)
    du_r = ref.FP[:n][fluid, du_c]
    du_g = after_momentum.float_properties_[:n][fluid, du_c]
    old_bar, floored = per_particle_bar(du_g, du_r, 1e-5 * abs(float(np.float32(ph.gy))))
    return {"rho": float(np.max(np.abs(rho_g - rho_r) / np.abs(rho_r))), "drho": drho,
            "du": bar_ratio(du_g, du_r, 1e-5 * a_m), "du_reordered": bar_ratio(du_rev, du_r, 1e-5 * a_m),
            "du_rel_floor": old_bar, "under_floor": floored, "entries": int(du_r.size),
            "du_max": float(np.max(np.abs(du_r))), "cancellation": float(np.max(a_m / (np.abs(du_r) + 1e-30)))}


def describe_pass(m):
    return (f"density {m['rho']:.3g} relative; dDensity {m['drho']:.3g} of the bar; acceleration "
            f"{m['du']:.3g} of the bar (reversed-order restatement: {m['du_reordered']:.3g}); against "
            f"1e-5 |du| + 1e-5 g instead: {m['du_rel_floor']:.3g} ({m['under_floor']} of {m['entries']} under "
            f"the floor); |du| <= {m['du_max']:.3g}, bound / |du| up to {m['cancellation']:.3g}")


def reference_momentum_magnitude(case):
    """sum |momentum terms| per particle and component (n, 2) of the first
    momentum pass from the case's tables (search, then sContinuity!)."""
    oph = oracle_physics(case.physics)
    ref = reference(case)
    assert ref.search(O.CRIT_SYMMETRY, 1, oph) == 0
    ref.selfaction(1, oph)
    return ref.term_magnitude(1, oph)


def flow_summary(case):
    """Largest speed, largest pressure and surge front (largest x of the fluid) of a case's tables."""
    h = case.host
    n = h.n_particles_
    fi = h.named_index_.float_table.index
    tag = h.int_properties_[:n, h.named_index_.int_table.index["Tag"] - 1]
    fluid = (h.is_alive_[:n] == 1) & (tag == 1)
    u = h.float_properties_[:n, fi["VelocityVec"] - 1: fi["VelocityVec"] + 1][fluid].astype(np.float64)
    return (float(np.max(np.hypot(u[:, 0], u[:, 1]))), float(np.max(h.float_properties_[:n, fi["Pressure"] - 1])),
            float(np.max(h.float_properties_[:n, fi["PositionVec"] - 1][fluid])))
