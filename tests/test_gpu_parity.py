"""libeppx (HIP, through the C ABI) against the restatement on the dam-break
cases.  EXACT math mode must agree bit for bit: same pair lists in the same
order, same floats.  FAST mode must agree within the tolerances stated below.
"""
import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from eppx import cases
    return cases.small(M=40)


def test_small_search_exact(parallel, small):
    """Pair lists (ids in list order) and every neighbour column of the first search."""
    ps, ns, ref = U.search_both(parallel, small)
    host = U.download(ps)
    assert int((host.is_alive_ == 0).sum()) == 153
    assert np.array_equal(host.cell_index_, ref.cell_index)
    U.assert_neighbour_columns_equal(host, ref)
    U.assert_state_equal(host, ref)
    off, js = U.export_pairs(ps, ns)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)


@pytest.mark.parametrize("route", ["step", "api"])
def test_small_steps_exact(parallel, small, route):
    """25 dam-break steps, via eppx_step and via the separate selfaction /
    interaction / search calls of main.jl:169-179."""
    from eppx import Algorithm, SPH
    ps, ns, ref = U.search_both(parallel, small)
    n = 25
    if route == "step":
        Algorithm.step(ps, ns, small.physics, n_steps=n)
    else:
        ph = small.physics
        for _ in range(n):
            Algorithm.selfaction(ps, SPH.sContinuity(ph), sync=False)
            Algorithm.interaction(ps, SPH.iMomentum(ph), sync=False)
            Algorithm.selfaction(ps, SPH.sMomentum(ph), sync=False)
            Algorithm.search(ps, small.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=SPH.sComputeKernelAndiContinuity(ph), sync=False)
        parallel.synchronize()
    err, pairs = ref.run_steps(n, U.oracle_physics(small.physics))
    assert err == 0
    host = U.download(ps)
    U.assert_state_equal(host, ref)
    U.assert_neighbour_columns_equal(host, ref)


def test_step_graph_replays_interleaved(parallel, small):
    """eppx_step replays a recorded hipGraph (one per buffer parity).  Replays
    interleaved with eager calls, host synchronisation and odd step counts
    must each leave the state the restatement has after as many steps."""
    from eppx import Algorithm, SPH
    ps, ns, ref = U.search_both(parallel, small)
    ph = small.physics
    oph = U.oracle_physics(ph)
    for seg, route in ((3, "step"), (2, "api"), (4, "step"), (1, "api"), (5, "step"), (2, "step")):
        if route == "step":
            Algorithm.step(ps, ns, ph, n_steps=seg)
        else:
            for _ in range(seg):
                Algorithm.selfaction(ps, SPH.sContinuity(ph), sync=False)
                Algorithm.interaction(ps, SPH.iMomentum(ph), sync=False)
                Algorithm.selfaction(ps, SPH.sMomentum(ph), sync=False)
                Algorithm.search(ps, small.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                                 action=SPH.sComputeKernelAndiContinuity(ph), sync=False)
            parallel.synchronize()
        err, _ = ref.run_steps(seg, oph)
        assert err == 0
        U.assert_state_equal(U.download(ps), ref,
                             names=("PositionVec", "VelocityVec", "Density", "Pressure", "dDensity", "Volume"))


def test_shepard_filter_exact(parallel, small):
    """iFilter + sFilter (main.jl:128-144) after a search, and eppx_step with
    the filter every 5 steps (main.jl:180-183): bit-identical to the restatement."""
    from eppx import Algorithm, SPH
    ph = small.physics
    oph = U.oracle_physics(ph)
    ps, ns, ref = U.search_both(parallel, small)
    Algorithm.interaction(ps, SPH.iFilter(ph))
    Algorithm.selfaction(ps, SPH.sFilter(ph))
    ref.interaction(2, oph)
    ref.selfaction(3, oph)
    names = ("Density", "Volume", "Pressure", "SumWeight", "SumWeightedDensity")
    U.assert_state_equal(U.download(ps), ref, names=names)
    Algorithm.step(ps, ns, ph, n_steps=12, filter_interval=5)
    err, _ = ref.run_steps(12, oph, filter_interval=5)
    assert err == 0
    U.assert_state_equal(U.download(ps), ref, names=names + ("PositionVec", "VelocityVec"))


def test_mass_changed_after_search(parallel, small):
    """Mass scaled between the search and the momentum pass: iMomentum! reads
    m_j at interaction time, so the pair terms must see the new masses."""
    from eppx import Algorithm, SPH
    ps, ns, ref = U.search_both(parallel, small)
    ph = small.physics
    Algorithm.selfaction(ps, SPH.ScaleField("Mass", 0, 1.5))
    Algorithm.interaction(ps, SPH.iMomentum(ph))
    ref.selfaction(10, column=ref.L.Mass, value=1.5)
    ref.interaction(1, U.oracle_physics(ph))
    U.assert_state_equal(U.download(ps), ref, names=("Mass", "dVelocityVec"))
    # the case started with one Mass and one H on every slot, which lets the cell-order
    # move skip both columns; the scaled masses (alive slots only) must move again
    Algorithm.step(ps, ns, ph, n_steps=3)
    err, _ = ref.run_steps(3, U.oracle_physics(ph))
    assert err == 0
    U.assert_state_equal(U.download(ps), ref, names=("Mass", "PositionVec", "VelocityVec", "Density"))


def test_small_1000_steps(parallel, small, oracle):
    """1000 steps (0.048 s of flow): bit-identical state; so the surge front
    (max x of fluid) is identical, well inside 0.5 dx; mass is exact."""
    from eppx import Algorithm
    ps, ns, ref = U.search_both(parallel, small)
    Algorithm.step(ps, ns, small.physics, n_steps=1000)
    err, _ = ref.run_steps(1000, U.oracle_physics(small.physics), threads=0)
    host = U.download(ps)
    U.assert_state_equal(host, ref, names=("PositionVec", "VelocityVec", "Density", "Pressure"))
    fluid = (ref.IP[: ref.n, ref.L.Tag] == 1) & (ref.is_alive[: ref.n] == 1)
    front_gpu = host.float_properties_[: ref.n, ref.L.PositionVec][fluid].max()
    front_ref = ref.FP[: ref.n, ref.L.PositionVec][fluid].max()
    assert abs(front_gpu - front_ref) <= 0.5 * small.dx
    m = ref.L.Mass
    assert host.float_properties_[: ref.n, m][fluid].astype(np.float64).sum() == \
        small.host.float_properties_[: ref.n, m][fluid].astype(np.float64).sum()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jittered_pairs_exact(parallel, seed):
    """Random sub-dx jitter of the fluid (ties broken generically): pair lists exact."""
    from eppx import cases
    case = cases.dam_break(0.01, 2.0, M=64)
    rng = np.random.default_rng(seed)
    h = case.host
    fluid = h.int_properties_[:, 0] == 1
    pos = h.float_properties_[:, 0:2]
    pos[fluid] += (rng.uniform(-0.45, 0.45, (fluid.sum(), 2)) * case.dx).astype(np.float32)
    ps, ns, ref = U.search_both(parallel, case)
    off, js = U.export_pairs(ps, ns)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    U.assert_state_equal(U.download(ps), ref, names=("dDensity",))


def test_fast_mode_within_tolerance(parallel, small):
    """FAST pair arithmetic on a developed flow (1000 steps in: the column
    collapsing, the cfg1 run's end state): one search, sContinuity! and
    iMomentum! against the exact restatement from the same tables, per
    particle.  density: |d rho| <= 1e-5 |rho|; dDensity and acceleration:
    within 1e-5 of the sum of |terms| of the particle's sum (the float32 bound
    of a sum taken in another order; du sums pressure terms that cancel down
    to about g, so 1e-5 |du| alone is below the rounding of its own terms)."""
    dev = U.developed(parallel, small, 1000)
    umax, pmax, front = U.flow_summary(dev)
    m = U.fast_momentum_pass(parallel, dev)
    print(f"FAST, small case 1000 steps in (|u| <= {umax:.3g} m/s, p <= {pmax:.3g} Pa, front {front:.3f} m): "
          + U.describe_pass(m))
    assert umax > 0.1 and front > 0.99  # really developed
    assert m["rho"] <= 1e-5
    assert m["drho"] <= 1.0
    assert m["du"] <= 1.0


def test_fast_mode_within_tolerance_wendland_mid_run(parallel):
    """The large benchmark's physics (Wendland C2, h = 2dx, one H: the
    hoisted FAST terms) at dx = 0.01 on a developed flow (3000 EXACT steps,
    0.096 s: hydrostatic pressure built up, the column collapsing): the
    bars of test_fast_mode_within_tolerance."""
    from eppx import SPH, cases
    case = cases.dam_break(0.01, 2.0, SPH.WendlandC2, M=64)
    dev = U.developed(parallel, case, 3000)
    umax, pmax, front = U.flow_summary(dev)
    m = U.fast_momentum_pass(parallel, dev)
    print(f"FAST, Wendland dx = 0.01, 3000 steps in (|u| <= {umax:.3g} m/s, p <= {pmax:.3g} Pa, "
          f"front {front:.3f} m): " + U.describe_pass(m))
    assert umax > 0.1
    assert m["rho"] <= 1e-5
    assert m["drho"] <= 1.0
    assert m["du"] <= 1.0


def test_large_fast_steps_within_tolerance(parallel):
    """The benchmark case at full size (2,030,018 particles): one call of 10
    fused FAST steps (the liveness skips and hoisted one-H terms included)
    against the same call in EXACT math, which the other tests hold
    bit-identical to the restatement.  The same particles alive, Mass exact,
    the same pair count (pair sets are exact in both modes while positions
    agree), density within 1e-5 relative, positions within 1e-6 m."""
    import ctypes as C
    from eppx import Algorithm, SPH, _lib, cases
    case = cases.large(M=64)
    L = _lib.lib()
    out = {}
    for math in ("exact", "fast"):
        parallel.set_math(math)
        try:
            ps, ns = U.device(parallel, case)
            Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=SPH.sComputeKernelAndiContinuity(case.physics))
            total = C.c_int64()
            _lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), parallel.ctx, "pair_total")
            Algorithm.step(ps, ns, case.physics, n_steps=10)
            _lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), parallel.ctx, "pair_total")
            host = U.download(ps)
            lay = ps.layout
            n = case.n_particles
            out[math] = (host.is_alive_[:n].copy(), host.float_properties_[:n, lay.Mass].copy(),
                         host.float_properties_[:n, lay.Density].astype(np.float64),
                         host.float_properties_[:n, lay.PositionVec: lay.PositionVec + 2].astype(np.float64),
                         total.value)
            del host
            ps.close()
            ns.close()
        finally:
            parallel.set_math("exact")
    (ae, me, re_, xe, pe), (af, mf, rf, xf, pf) = out["exact"], out["fast"]
    assert np.array_equal(ae, af) and (ae == 1).all()
    assert np.array_equal(me, mf)
    assert pe == pf and pe > 0
    alive = ae == 1
    rho_err = np.max(np.abs(rf[alive] - re_[alive]) / np.abs(re_[alive]))
    pos_err = np.max(np.abs(xf[alive] - xe[alive]))
    print(f"2M case, 10 steps FAST vs EXACT: density {rho_err:.3g} relative, positions {pos_err:.3g} m, "
          f"{pe} pairs over the steps")
    assert rho_err <= 1e-5
    assert pos_err <= 1e-6


def test_large_first_search_exact(parallel):
    """2,030,018 particles (dx = 0.001): every pair list and drho identical."""
    from eppx import cases
    case = cases.large(M=64)
    ps, ns, ref = U.search_both(parallel, case)
    off, js = U.export_pairs(ps, ns)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert off[-1] == roff[-1]
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    U.assert_state_equal(U.download(ps), ref, names=("dDensity",))


# ------------------------------------------------------------- error paths

def test_capacity_overflow_reported(parallel, small):
    """max_neighbour_number below the real neighbour count: the reference writes
    out of bounds; libeppx reports EPPX_E_CAPACITY."""
    from eppx import Algorithm, EppxError, SPH
    ps, ns = U.device(parallel, small, M=8)
    with pytest.raises(EppxError) as e:
        Algorithm.search(ps, small.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(small.physics))
    assert e.value.status == 3
    parallel.synchronize()  # the flag is cleared once reported


def test_outside_particle_dies_and_stays_dead(parallel, small):
    from eppx import Algorithm, SPH
    case = small
    h = case.host
    ps, ns = U.device(parallel, case)
    host = U.download(ps)
    host.float_properties_[7, 0] = 100.0  # far outside the domain
    from eppx import Class
    Class.syncto(ps, host)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    Algorithm.step(ps, ns, case.physics, n_steps=3)
    out = U.download(ps)
    assert out.is_alive_[7] == 0 and out.cell_index_[7] == 0
    assert int((out.is_alive_ == 0).sum()) == 154
    assert ps.get_alive_n_particles() == h.n_particles_ - 154


def test_interaction_before_search_is_rejected(parallel, small):
    from eppx import Algorithm, EppxError, SPH
    ps, ns = U.device(parallel, small)
    ps.last_ns = ns
    with pytest.raises(EppxError) as e:
        Algorithm.interaction(ps, SPH.iMomentum(small.physics))
    assert e.value.status == 5


def test_zero_particles(parallel, small):
    from eppx import Algorithm, Class, SPH
    ps = Class.ParticleSystem(parallel, small.domain, 0, small.int_nt, small.float_nt, small.basic_parameters,
                              n_capacity=16)
    ns = Class.NeighbourSystem(parallel, small.domain, [(1, 1), (1, 2), (2, 1)], max_neighbour_number=40)
    Algorithm.search(ps, small.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(small.physics))
    Algorithm.step(ps, ns, small.physics, n_steps=2)
    off, js = U.export_pairs(ps, ns)
    assert off.tolist() == [0] and js.size == 0


def test_steps_split_over_calls_exact(parallel, small):
    """The same 12 steps (filter every 4) in calls of 1, 3, 1 and 7 steps: a
    call's first step applies sContinuity! itself, later ones find it applied
    by the previous search, the last leaves it pending -- bit-identical to the
    restatement, with the same pairs."""
    from eppx import Algorithm
    ph = small.physics
    oph = U.oracle_physics(ph)
    ps, ns, ref = U.search_both(parallel, small)
    done = 0
    for n in (1, 3, 1, 7):
        Algorithm.step(ps, ns, ph, n_steps=n, filter_interval=4, step_offset=done)
        done += n
    err, pairs = ref.run_steps(12, oph, filter_interval=4)
    assert err == 0
    host = U.download(ps)
    U.assert_state_equal(host, ref, names=U.STATE_FLOAT + ("SumWeight", "SumWeightedDensity"))
    U.assert_neighbour_columns_equal(host, ref)
    off, _ = U.export_pairs(ps, ns)
    assert off[-1] == pairs[-1]


def test_particles_dying_inside_a_call_exact(parallel):
    """Fluid particles thrown out of the domain die in the middle of a
    10-step call (the later steps leave dDensity of the alive slots alone):
    every particle's state, the dead ones' included, is bit-identical to the
    restatement, and the pairs are the same."""
    from eppx import Algorithm, cases
    case = cases.small(M=64)
    h = case.host
    L = h.named_index_.float_table.index
    fluid = np.flatnonzero(h.int_properties_[: h.n_particles_, h.named_index_.int_table.index["Tag"] - 1] == 1)
    u = L["VelocityVec"] - 1
    throw = fluid[::700][:6]  # a few fluid particles across the column
    h.float_properties_[throw, u] = 2.0e4  # m/s: about 1 m per step, out of the tank within a few steps
    ph = case.physics
    oph = U.oracle_physics(ph)
    ps, ns, ref = U.search_both(parallel, case, M=64)  # a thrown particle may land among others
    alive0 = ps.get_alive_n_particles()
    Algorithm.step(ps, ns, ph, n_steps=10)
    err, pairs = ref.run_steps(10, oph)
    assert err == 0
    host = U.download(ps)
    assert ps.get_alive_n_particles() < alive0  # some died during the call
    U.assert_state_equal(host, ref, alive_only=False)
    U.assert_neighbour_columns_equal(host, ref)
    off, _ = U.export_pairs(ps, ns)
    assert off[-1] == pairs[-1]


@pytest.mark.parametrize("start", ["walls_moving_dvel", "momentum_before_step"])
def test_step_calls_with_nonzero_dvelocity_exact(parallel, start):
    """A step call whose first step finds non-zero dVelocity: on wall particles
    (the momentum pass never writes theirs, so it stays and the permutes must
    move it in every step), or on the fluid after a lone iMomentum! (the first
    step adds to it).  6 steps, bit-identical to the restatement."""
    from eppx import Algorithm, SPH, cases
    case = cases.small(M=40)
    h = case.host
    ph = case.physics
    oph = U.oracle_physics(ph)
    if start == "walls_moving_dvel":
        tag = h.int_properties_[: h.n_particles_, h.named_index_.int_table.index["Tag"] - 1]
        du = h.named_index_.float_table.index["dVelocityVec"] - 1
        walls = np.flatnonzero(tag == 2)[::50]
        h.float_properties_[walls, du] = 0.25
        h.float_properties_[walls, du + 1] = -0.5
    ps, ns, ref = U.search_both(parallel, case)
    if start == "momentum_before_step":
        Algorithm.interaction(ps, SPH.iMomentum(ph))
        ref.interaction(1, oph)
    for n in (1, 6):  # a one-step call, then one whose later steps may skip dVelocity
        Algorithm.step(ps, ns, ph, n_steps=n)
        err, _ = ref.run_steps(n, oph)
        assert err == 0
        U.assert_state_equal(U.download(ps), ref, alive_only=False)


def _fast_steps_errors(parallel, case, n):
    """n fused FAST steps on the device and n steps of the restatement from the
    same tables.  Per particle: density relative; velocity in units of
    1e-5 |u| + 1e-5 t (A + g), where A = sum |momentum terms| at the start
    (the per-step acceleration bar, integrated over the run's time t);
    positions in float32 spacings."""
    from eppx import Algorithm
    a_m = U.reference_momentum_magnitude(case)
    parallel.set_math("fast")
    try:
        ps, ns, ref = U.search_both(parallel, case)
        Algorithm.step(ps, ns, case.physics, n_steps=n)
        host = U.download(ps)
    finally:
        parallel.set_math("exact")
    err, _ = ref.run_steps(n, U.oracle_physics(case.physics))
    assert err == 0
    alive = (ref.is_alive[: ref.n] == 1) & (host.is_alive_[: ref.n] == 1)
    L = ref.L
    g = host.float_properties_[: ref.n][alive].astype(np.float64)
    r = ref.FP[: ref.n][alive].astype(np.float64)
    rho = np.max(np.abs(g[:, L.Density] - r[:, L.Density]) / np.abs(r[:, L.Density]))
    xr = r[:, L.PositionVec: L.PositionVec + 2]
    pos = np.max(np.abs(g[:, L.PositionVec: L.PositionVec + 2] - xr) / np.spacing(np.abs(xr).astype(np.float32)))
    t = n * float(np.float32(case.physics.dt))
    vr = r[:, L.VelocityVec: L.VelocityVec + 2]
    vel = U.bar_ratio(g[:, L.VelocityVec: L.VelocityVec + 2], vr, 1e-5 * np.abs(vr) + 1e-5 * t * (a_m[alive] + 9.81))
    return rho, pos, vel, int((~alive).sum() - (ref.is_alive[: ref.n] != 1).sum())


@pytest.mark.parametrize("vary_h,start", [(False, "rest"), (True, "rest"), (False, "developed")])
def test_fast_steps_within_tolerance(parallel, vary_h, start):
    """20 steps of the fused FAST step (hinv hoisted for one H, or per pair
    for varying H), from rest or from the flow 1000 steps in, against the
    exact restatement from the same tables, per particle: density within 1e-5
    relative; velocity within 1e-5 relative plus the per-step acceleration
    bar integrated over the run (1e-5 t (sum |terms| + g)); positions within
    one float32 spacing.  No particle dies in one and not the other."""
    from eppx import cases
    case = cases.small(M=64)
    if vary_h:
        rng = np.random.default_rng(21)
        h = case.host
        c = h.named_index_.float_table.index["H"] - 1
        h.float_properties_[: h.n_particles_, c] *= rng.uniform(0.8, 1.0, h.n_particles_).astype(np.float32)
    if start == "developed":
        case = U.developed(parallel, case, 1000)
    rho, pos, vel, extra_dead = _fast_steps_errors(parallel, case, 20)
    print(f"fast vs exact after 20 steps from {start}: rho {rho:.3g}, pos {pos:.3g} spacings, vel {vel:.3g} of "
          f"the bar")
    assert extra_dead == 0
    assert rho <= 1e-5
    assert pos <= 1.0
    assert vel <= 1.0


def test_fast_1000_steps_surge_front(parallel, small):
    """The trajectory bar: 1000 fused FAST steps of the small dam break (0.048 s,
    the column collapsing) against the EXACT run, which the other tests hold
    bit-identical to the restatement: the surge front (largest x of the fluid)
    within 0.5 dx, the same particles alive, Mass exact, density within 1e-5
    relative per particle."""
    from eppx import Algorithm
    out = {}
    for math in ("exact", "fast"):
        parallel.set_math(math)
        try:
            ps, ns = U.device(parallel, small)
            Algorithm.search(ps, small.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=__import__("eppx").SPH.sComputeKernelAndiContinuity(small.physics))
            Algorithm.step(ps, ns, small.physics, n_steps=1000)
            out[math] = U.download(ps)
            ps.close()
            ns.close()
        finally:
            parallel.set_math("exact")
    e, f = out["exact"], out["fast"]
    n = small.n_particles
    fi = e.named_index_.float_table.index
    tag = e.int_properties_[:n, e.named_index_.int_table.index["Tag"] - 1]
    assert np.array_equal(e.is_alive_[:n], f.is_alive_[:n])
    fluid = (e.is_alive_[:n] == 1) & (tag == 1)
    x = fi["PositionVec"] - 1
    front_e = float(e.float_properties_[:n, x][fluid].max())
    front_f = float(f.float_properties_[:n, x][fluid].max())
    m = fi["Mass"] - 1
    assert np.array_equal(e.float_properties_[:n, m], f.float_properties_[:n, m])
    d = fi["Density"] - 1
    alive = e.is_alive_[:n] == 1
    rho = float(np.max(np.abs(f.float_properties_[:n, d][alive].astype(np.float64) - e.float_properties_[:n, d][alive])
                       / e.float_properties_[:n, d][alive]))
    print(f"1000 FAST steps: surge front {front_f:.6f} m vs {front_e:.6f} m exact "
          f"({abs(front_f - front_e) / small.dx:.3g} dx), density {rho:.3g} relative")
    assert abs(front_f - front_e) <= 0.5 * small.dx
    assert front_e > 0.99 + 0.2 * small.dx  # the front moved (it starts at the last lattice column, 0.99 m)
    assert rho <= 1e-5
