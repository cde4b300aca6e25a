"""The reference example's own physics on the device, against the restatement:
example/collapse_dry/same/main.jl:57-70 with casegen.jl's geometry (dx = 0.01,
cubic spline h = 1.5 dx, cell gap 3 dx, max_neighbour_number 40): the linear
EOS p = c0^2 (rho - rho0) + p0 with c0 = 120, iClassicViscosity! with
mu0 = 1e-3 (main.jl:110,114 -> Viscosity.jl:10-26), g = 9.8,
dt = 0.2 h0 / c0, and the Shepard filter every 20 steps (main.jl:180-183).
EXACT math: bit-identical.  FAST math: per particle, density within 1e-5
relative; dDensity and acceleration within 1e-5 of the sum of |terms| of
the particle's sum (the float32 bound of a sum taken in another order)."""
import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu

G = 9.8


@pytest.fixture(scope="module")
def example():
    from eppx import SPH, cases
    case = cases.dam_break(0.01, 1.5, SPH.CubicSpline, M=40)
    case.physics = SPH.example_physics(0.01)
    return case


def test_example_physics_constants(example):
    from eppx import _lib
    ph = example.physics
    assert ph.eos == _lib.EOS_LINEAR and ph.visc == _lib.VISC_CLASSIC and ph.kernel == _lib.KERNEL_CUBIC
    assert np.float32(ph.c0) == np.float32(120.0) and np.float32(ph.gy) == np.float32(-9.8)
    assert np.float32(ph.dt) == np.float32(np.float32(0.1 * np.float64(np.float32(0.015)) / 120.0) * 2)
    assert example.n_particles == 20000 + 2 * 3 * 300 + 3 * 406


@pytest.mark.parametrize("route", ["step", "api"])
def test_example_physics_steps_exact(parallel, example, route):
    """60 steps of main.jl's loop (filter at 20, 40, 60), through eppx_step and
    through the separate calls of main.jl:169-183: bit-identical state and
    neighbour lists."""
    from eppx import Algorithm, SPH
    ph = example.physics
    oph = U.oracle_physics(ph)
    ps, ns, ref = U.search_both(parallel, example)
    n = 60
    if route == "step":
        Algorithm.step(ps, ns, ph, n_steps=n, filter_interval=20)
    else:
        for k in range(1, n + 1):
            Algorithm.selfaction(ps, SPH.sContinuity(ph), sync=False)
            Algorithm.interaction(ps, SPH.iMomentum(ph), sync=False)
            Algorithm.selfaction(ps, SPH.sMomentum(ph), sync=False)
            Algorithm.search(ps, example.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=SPH.sComputeKernelAndiContinuity(ph), sync=False)
            if k % 20 == 0:
                Algorithm.interaction(ps, SPH.iFilter(ph), sync=False)
                Algorithm.selfaction(ps, SPH.sFilter(ph), sync=False)
        parallel.synchronize()
    err, pairs = ref.run_steps(n, oph, filter_interval=20)
    assert err == 0
    host = U.download(ps)
    U.assert_state_equal(host, ref, names=U.STATE_FLOAT + ("SumWeight", "SumWeightedDensity"), alive_only=False)
    U.assert_neighbour_columns_equal(host, ref)
    off, _ = U.export_pairs(ps, ns)
    assert off[-1] == pairs[-1]
    # the linear EOS really ran: p = c0^2 (rho - rho0) + p0 on every alive particle
    L = ref.L
    alive = ref.is_alive[: ref.n] == 1
    rho = ref.FP[: ref.n][alive, L.Density]
    c02 = np.float32(ph.c0) * np.float32(ph.c0)
    assert np.array_equal(ref.FP[: ref.n][alive, L.Pressure], c02 * (rho - np.float32(1000.0)) + np.float32(0.0))


def test_example_physics_fast_within_tolerance(parallel, example):
    """A developed flow (4000 EXACT steps, 0.1 s, filter every 20), then one
    search, sContinuity! and iMomentum! in FAST math against the restatement
    from the same tables: density within 1e-5 relative per particle; dDensity
    and acceleration within 1e-5 of the sum of |terms| of each particle's sum
    (the float32 bound of a sum taken in another order)."""
    dev = U.developed(parallel, example, 4000, filter_interval=20)
    umax, pmax, front = U.flow_summary(dev)
    m = U.fast_momentum_pass(parallel, dev)
    print(f"example physics, FAST 4000 steps in (|u| <= {umax:.3g} m/s, p <= {pmax:.3g} Pa, front {front:.3f} m): "
          + U.describe_pass(m))
    assert umax > 0.1
    assert m["rho"] <= 1e-5
    assert m["drho"] <= 1.0
    assert m["du"] <= 1.0


def test_example_physics_fast_steps_within_tolerance(parallel, example):
    """20 fused FAST steps from rest against the exact restatement, per
    particle: density within 1e-5 relative; velocity (the acceleration
    integrated over the steps) within 1e-5 relative plus 1e-5 g t; positions
    within one float32 spacing of the restatement's."""
    from eppx import Algorithm
    ph = example.physics
    parallel.set_math("fast")
    try:
        ps, ns, ref = U.search_both(parallel, example)
        Algorithm.step(ps, ns, ph, n_steps=20)
        host = U.download(ps)
    finally:
        parallel.set_math("exact")
    err, _ = ref.run_steps(20, U.oracle_physics(ph))
    assert err == 0
    assert np.array_equal(host.is_alive_[: ref.n], ref.is_alive[: ref.n])
    L = ref.L
    alive = ref.is_alive[: ref.n] == 1
    g, r = host.float_properties_[: ref.n][alive], ref.FP[: ref.n][alive]
    rho = float(np.max(np.abs(g[:, L.Density].astype(np.float64) - r[:, L.Density]) / np.abs(r[:, L.Density])))
    t = 20 * float(np.float32(ph.dt))
    vel, floored = U.per_particle_bar(g[:, L.VelocityVec: L.VelocityVec + 2], r[:, L.VelocityVec: L.VelocityVec + 2],
                                      1e-5 * G * t)
    xr = r[:, L.PositionVec: L.PositionVec + 2]
    pos = float(np.max(np.abs(g[:, L.PositionVec: L.PositionVec + 2].astype(np.float64) - xr) /
                       np.spacing(np.abs(xr))))
    print(f"example physics, 20 FAST steps: density {rho:.3g}, velocity {vel:.3g} of the bar ({floored} under "
          f"the floor), positions {pos:.3g} spacings")
    assert rho <= 1e-5
    assert vel <= 1.0
    assert pos <= 1.0
