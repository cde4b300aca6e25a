"""The benchmark's own configuration against the restatement: the
2,030,018-particle dam break (dx = 0.001, Wendland C2, h = 2dx, Tait EOS,
artificial viscosity; main.jl:166-179 with Pressure.jl:24-60 and
Viscosity.jl:28-52).

* FAST math with the autotuned pair shapes (what bench.py times), on the flow
  10,000 steps in (both sides start from the same tables): the acceleration
  of a momentum pass per particle, then two fused steps (the bench's
  kernels: continuity + EOS, momentum + move + cell keys, the search with
  the next continuity fused, momentum + move, the plain search).
  Bars, per particle: density 1e-5 relative; dDensity and acceleration
  within 1e-5 of the sum of |terms| of the particle's sum (the float32 bound
  of a sum taken in another order: du sums pressure terms of ~1e3 m/s^2 that
  cancel down to about g, so 1e-5 |du| is below the rounding of its own
  terms); after the fused steps, density 1e-5 relative, positions within one
  float32 spacing, velocity within twice the per-step bound integrated over
  the steps (the second step starts from inputs that differ at the ulp
  level).  Pair sets bit-exact, Mass exact.
* EXACT math: three steps bit-identical to the restatement, pairs included.
"""
import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu

G = 9.81


@pytest.fixture(scope="module")
def large():
    from eppx import cases
    return cases.large(M=64)


def test_large_fast_matches_restatement(parallel, large):
    """From the flow 10,000 FAST steps in (32 ms: the pressure wave has crossed
    the column, pressure is hydrostatic, the column has started to collapse),
    the restatement and the device start from the same tables."""
    from eppx import Algorithm, SPH
    ph = large.physics
    oph = U.oracle_physics(ph)
    dev = U.developed(parallel, large, 10000, math="fast")
    umax, pmax, front = U.flow_summary(dev)
    print(f"2M flow 10,000 steps in: |u| <= {umax:.3g} m/s, p <= {pmax:.3g} Pa, front {front:.4f} m")
    assert umax > 0.05 and pmax > 1.5e4
    parallel.set_math("fast")
    try:
        ps, ns = U.device(parallel, dev)
        Algorithm.search(ps, dev.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(ph))
        tune = Algorithm.autotune(ps, ns, ph, trials=1)
        ps.close()
        ns.close()
    finally:
        parallel.set_math("exact")
    shapes = {"search": (tune["search"]["block"], tune["search"]["staged"]),
              "momentum": (tune["momentum"]["block"], tune["momentum"]["staged"])}
    print(f"autotuned shapes: {shapes}")

    # (1) one search + sContinuity! + iMomentum!, per particle
    m = U.fast_momentum_pass(parallel, dev, shapes)
    print("2M FAST search + momentum pass: " + U.describe_pass(m))
    assert m["rho"] <= 1e-5
    assert m["drho"] <= 1.0
    assert m["du"] <= 1.0
    a_m = U.reference_momentum_magnitude(dev)

    # (2) two fused steps, the benchmark's kernels
    parallel.set_math("fast")
    try:
        ps, ns, ref = U.search_both(parallel, dev)
        Algorithm.set_pair_shape(ps, search=shapes["search"], momentum=shapes["momentum"])
        Algorithm.step(ps, ns, ph, n_steps=2)
        host = U.download(ps)
        off, js = U.export_pairs(ps, ns)
        ps.close()
        ns.close()
    finally:
        parallel.set_math("exact")
    err, pairs = ref.run_steps(2, oph)
    assert err == 0
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    L = ref.L
    n = ref.n
    assert np.array_equal(host.is_alive_[:n], ref.is_alive[:n])
    assert np.array_equal(host.float_properties_[:n, L.Mass], ref.FP[:n, L.Mass])
    alive = ref.is_alive[:n] == 1
    g, r = host.float_properties_[:n][alive], ref.FP[:n][alive]
    rho = float(np.max(np.abs(g[:, L.Density].astype(np.float64) - r[:, L.Density]) / np.abs(r[:, L.Density])))
    xr = r[:, L.PositionVec: L.PositionVec + 2].astype(np.float64)
    dpos = np.abs(g[:, L.PositionVec: L.PositionVec + 2] - xr)
    pos = float(np.max(dpos / np.spacing(np.abs(xr).astype(np.float32))))
    # After the first FAST step the second one starts from inputs that differ at the ulp level (velocities
    # by ~1e-7 m/s, a position by one spacing), and where the flow moves almost rigidly the relative
    # velocities of neighbours are ~1e-4 of |u|: the terms of the second step amplify those input
    # differences beyond the per-step bound, which part (1) holds from identical inputs.  So velocity is
    # held to twice the per-step bound integrated over the steps, and dDensity (the pending sum of the
    # last search, the most input-sensitive) is reported only.
    t = 2 * float(np.float32(ph.dt))
    vr = r[:, L.VelocityVec: L.VelocityVec + 2].astype(np.float64)
    vbar = 1e-5 * np.abs(vr) + 1e-5 * t * (a_m[alive] + G)
    vel = U.bar_ratio(g[:, L.VelocityVec: L.VelocityVec + 2], vr, vbar)
    over = int((np.abs(g[:, L.VelocityVec: L.VelocityVec + 2] - vr) > vbar).any(axis=1).sum())
    a_c = ref.term_magnitude(0, oph)[alive]
    drho = U.bar_ratio(g[:, L.dDensity], r[:, L.dDensity], 1e-5 * a_c)
    print(f"2M, 2 fused FAST steps: density {rho:.3g} relative; positions {pos:.3g} spacings "
          f"({int((dpos > 0).any(axis=1).sum())} particles differ); velocity {vel:.3g} of 1e-5 |u| + 1e-5 t "
          f"(bound + g) ({over} of {int(alive.sum())} particles above 1); dDensity {drho:.3g} of 1e-5 x its bound; "
          f"{int(pairs.sum())} pairs")
    assert rho <= 1e-5
    assert pos <= 1.0
    assert vel <= 2.0


def test_large_exact_steps_bit_identical(parallel, large):
    """Three EXACT steps at 2,030,018 particles: every particle's state and the
    pair lists identical to the restatement's."""
    from eppx import Algorithm
    ps, ns, ref = U.search_both(parallel, large)
    Algorithm.step(ps, ns, large.physics, n_steps=3)
    host = U.download(ps)
    off, js = U.export_pairs(ps, ns)
    ps.close()
    ns.close()
    err, pairs = ref.run_steps(3, U.oracle_physics(large.physics))
    assert err == 0
    U.assert_state_equal(host, ref, alive_only=False)
    roff, rjs = U.reference_pairs_in_order(ref)
    assert np.array_equal(off, roff) and np.array_equal(js, rjs)
    assert off[-1] == pairs[-1]
