"""Pins the oracle where the reference's tests say nothing (physics values):
closed forms, invariants and brute force.  Each check is paired with a
restatement variant that plants one plausible mistake (OR_MUTATION in
oracle/eppx_oracle.c) and must make that check fail."""
import math

import numpy as np
import pytest

M = 64


def lattice(O, nx, ny, dx, h_factor=2.0, tags=None, variant=None, vel=None, rho=None, p=None, jitter=0.0, seed=0):
    """nx x ny fluid lattice in the dam-break layout, domain cell = 2h."""
    int_nt, float_nt = O.dam_break_layout(M)
    n = nx * ny
    L = O.make_layout(int_nt, float_nt, n)
    IP = np.zeros((n, L.n_int), np.int32, order="F")
    FP = np.zeros((n, L.n_float), np.float32, order="F")
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = (0.5 + ii.ravel()) * dx
    y = (0.5 + jj.ravel()) * dx
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter, jitter, n) * dx
        y = y + rng.uniform(-jitter, jitter, n) * dx
    x, y = x.astype(np.float32), y.astype(np.float32)
    FP[:, L.PositionVec], FP[:, L.PositionVec + 1] = x, y
    IP[:, L.Tag] = 1 if tags is None else tags
    IP[:, L.IsMovable] = 1
    vol = np.float32(dx * dx)
    FP[:, L.Volume] = vol
    FP[:, L.Density] = 1000.0 if rho is None else rho(x, y)
    FP[:, L.Mass] = FP[:, L.Density] * vol  # consistent volumes V = m / rho = dx^2
    FP[:, L.H] = np.float32(h_factor * dx)
    if vel is not None:
        u, v = vel(x, y)
        FP[:, L.VelocityVec], FP[:, L.VelocityVec + 1] = u, v
    if p is not None:
        FP[:, L.Pressure] = p(x, y)
    gap = 2 * h_factor * dx
    d = O.domain2d(gap, -gap, -gap, nx * dx + gap, ny * dx + gap)
    adj = np.array([[1, 1], [1, 0]], np.int32)
    return O.System(IP, FP, np.ones(n, np.int32), np.zeros(n, np.int32), n, L, d, adj, M, variant=variant), x, y


def phys(O, kernel=1, **kw):
    base = dict(fluid_tag=1, wall_tag=2, kernel=kernel, eos=1, visc=1, dt=1e-4, rho0=1000.0, p0=0.0,
                c0=62.6418, tait_b=1000.0 * 62.6418 ** 2 / 7, gx=0.0, gy=-9.81, alpha=0.1, beta=0.1,
                av_c=62.6418, mu=1e-3, coefficient=0.0)
    base.update(kw)
    return O.physics(**base)


def interior(x, y, nx, ny, dx, margin=5):
    return (x > margin * dx) & (x < (nx - margin) * dx) & (y > margin * dx) & (y < (ny - margin) * dx)


# ------------------------------------------------------------------ kernels

@pytest.mark.parametrize("kernel", [0, 1])
def test_kernel_normalisation_support_and_centre(oracle, kernel):
    lib = oracle.lib()
    h = 0.01
    hinv = 1.0 / h
    r = np.linspace(0, 2 * h, 20001)
    w = np.array([lib.or_kernel_value(kernel, float(x), hinv) for x in r])
    integral = np.trapezoid(2 * math.pi * r * w, r)
    assert integral == pytest.approx(1.0, rel=2e-4)                 # ∫ W dA = 1
    assert lib.or_kernel_value(kernel, 2 * h, hinv) == 0.0           # compact support
    assert lib.or_kernel_gradient(kernel, 2 * h * 1.0001, hinv) == 0.0
    sigma = 10 / 7 / math.pi if kernel == 0 else 7 / 4 / math.pi
    assert lib.or_kernel_value(kernel, 0.0, hinv) == pytest.approx(sigma / h ** 2, rel=1e-6)
    assert lib.or_kernel_value0(kernel, h) == pytest.approx(sigma / h ** 2, rel=1e-6)


def gradient_matches_derivative(lib, kernel):
    h = 0.01
    hinv = 1.0 / h
    for q in (0.2, 0.7, 0.99, 1.01, 1.3, 1.8):
        r = q * h
        e = 1e-4 * h
        fd = (lib.or_kernel_value(kernel, r + e, hinv) - lib.or_kernel_value(kernel, r - e, hinv)) / (2 * e)
        if abs(lib.or_kernel_gradient(kernel, r, hinv) - fd) > 2e-3 * abs(fd):
            return False
    return True


@pytest.mark.parametrize("kernel", [0, 1])
def test_gradient_is_derivative(oracle, kernel):
    assert gradient_matches_derivative(oracle.lib(), kernel)


def test_gradient_check_catches_dropped_factor(oracle):
    assert not gradient_matches_derivative(oracle.mutant(5), 1)


def test_means(oracle):
    lib = oracle.lib()
    assert lib.or_harmonic(0.5, 0.5) == 0.5
    assert lib.or_harmonic(1.0, 3.0) == pytest.approx(1.5)
    assert lib.or_invharmonic(1.0, 3.0) == pytest.approx(1 / 1.5)


# -------------------------------------------------------------- continuity

def divergence_error(O, variant=None, tags=None):
    """u = (a x, 0) on a lattice: interior drho ≈ -rho0 * a (SPH consistency)."""
    nx = ny = 30
    dx, a = 0.01, 0.5
    s, x, y = lattice(O, nx, ny, dx, variant=variant, tags=tags, vel=lambda x, y: (a * x, 0 * y))
    assert s.search(O.CRIT_SYMMETRY, 1, phys(O)) == 0
    drho = s.column("dDensity")
    sel = interior(x, y, nx, ny, dx)
    if tags is not None:
        sel &= tags == 1
    return np.max(np.abs(drho[sel] - (-1000.0 * a)) / (1000.0 * a))


def test_continuity_divergence(oracle):
    assert divergence_error(oracle) < 0.03


def test_continuity_balanced_equals_classic_for_uniform_density(oracle):
    """Fluid-wall (balanced) terms equal fluid-fluid terms when m = rho V."""
    tags = np.where((np.arange(900) % 2) == 0, 1, 2).astype(np.int32)
    assert divergence_error(oracle, tags=tags) < 0.03


def test_continuity_check_catches_sign(oracle):
    assert divergence_error(oracle, variant=oracle.mutant(1)) > 1.0


# ---------------------------------------------------------------- momentum

def pair_system(O, variant=None, rho=(1000.0, 1200.0), p=(5e4, 2e4), u=((0.3, 0.1), (-0.2, 0.05))):
    """Two fluid particles inside each other's support."""
    int_nt, float_nt = O.dam_break_layout(M)
    L = O.make_layout(int_nt, float_nt, 2)
    IP = np.zeros((2, L.n_int), np.int32, order="F")
    FP = np.zeros((2, L.n_float), np.float32, order="F")
    IP[:, L.Tag] = 1
    IP[:, L.IsMovable] = 1
    FP[:, L.PositionVec] = [0.100, 0.112]
    FP[:, L.PositionVec + 1] = [0.100, 0.095]
    FP[:, L.VelocityVec] = [u[0][0], u[1][0]]
    FP[:, L.VelocityVec + 1] = [u[0][1], u[1][1]]
    FP[:, L.Density] = rho
    FP[:, L.Pressure] = p
    FP[:, L.Mass] = [0.1, 0.13]
    FP[:, L.Volume] = FP[:, L.Mass] / FP[:, L.Density]
    FP[:, L.H] = [0.010, 0.011]
    d = O.domain2d(0.05, 0.0, 0.0, 0.3, 0.3)
    s = O.System(IP, FP, np.ones(2, np.int32), np.zeros(2, np.int32), 2, L, d, np.array([[1, 1], [1, 0]]), M,
                 variant=variant)
    return s


def momentum_residual(O, variant=None, **kw):
    s = pair_system(O, variant, **kw)
    ph = phys(O)
    assert s.search(O.CRIT_SYMMETRY, 1, ph) == 0
    assert s.count().tolist() == [1, 1]
    s.interaction(1, ph)
    du = s.FP[:2, s.L.dVelocityVec: s.L.dVelocityVec + 2].astype(np.float64)
    m = s.FP[:2, s.L.Mass].astype(np.float64)
    total = m[0] * du[0] + m[1] * du[1]
    return np.max(np.abs(total)) / np.max(np.abs(m[:, None] * du)), du


def test_momentum_antisymmetric(oracle):
    """Pairwise momentum exchange of pressure + artificial viscosity cancels."""
    res, du = momentum_residual(oracle)
    assert np.max(np.abs(du)) > 1.0 and res < 1e-5


def test_antisymmetry_catches_transposed_density(oracle):
    assert momentum_residual(oracle, oracle.mutant(2))[0] > 1e-2


def viscosity_acts(O, variant=None):
    """p = 0: artificial viscosity alone must slow an approaching pair and leave a
    receding pair untouched (Viscosity.jl:28-52)."""
    approach = ((0.3, 0.0), (-0.3, 0.0))
    recede = ((-0.3, 0.0), (0.3, 0.0))
    out = []
    for u in (approach, recede):
        s = pair_system(O, variant, p=(0.0, 0.0), u=u)
        ph = phys(O)
        s.search(O.CRIT_SYMMETRY, 1, ph)
        s.interaction(1, ph)
        du = s.FP[:2, s.L.dVelocityVec: s.L.dVelocityVec + 2].astype(np.float64)
        rel_u = np.array(u[0]) - np.array(u[1])
        out.append(float(np.dot(du[0] - du[1], rel_u)))
    return out[0] < 0.0 and out[1] == 0.0


def test_artificial_viscosity_dissipates(oracle):
    assert viscosity_acts(oracle)


@pytest.mark.parametrize("mutation", [4, 6])
def test_viscosity_check_catches_dropped_or_flipped_term(oracle, mutation):
    assert not viscosity_acts(oracle, oracle.mutant(mutation))


def test_pressure_gradient(oracle):
    """p = P0 + k y with rho = rho0 (1 + e y): interior du_y ≈ -(dp/dy) / rho."""
    nx = ny = 30
    dx, k = 0.01, 2e4
    s, x, y = lattice(oracle, nx, ny, dx, rho=lambda x, y: 1000.0 * (1 + 0.5 * y), p=lambda x, y: 1e5 + k * y)
    ph = phys(oracle)
    s.search(oracle.CRIT_SYMMETRY, 1, ph)
    s.interaction(1, ph)
    du_y = s.FP[: s.n, s.L.dVelocityVec + 1]
    sel = interior(x, y, nx, ny, dx)
    expect = -k / (1000.0 * (1 + 0.5 * y[sel]))
    assert np.max(np.abs(du_y[sel] - expect) / np.abs(expect)) < 0.05


# -------------------------------------------------------------- self actions

def free_fall_error(O, variant=None, steps=20):
    s = pair_system(O, variant, u=((0.0, 0.0), (0.0, 0.0)))
    s.FP[1, s.L.PositionVec] = 0.25  # far from particle 0: no neighbours
    dt = np.float32(1e-3)
    ph = phys(O, dt=dt, gy=-9.81)
    y0 = float(s.FP[0, s.L.PositionVec + 1])
    for _ in range(steps):
        s.selfaction(2, ph)
    t = steps * float(dt)
    exact = y0 - 0.5 * 9.81 * t * t
    return abs(float(s.FP[0, s.L.PositionVec + 1]) - exact) / (0.5 * 9.81 * t * t)


def test_integrator_free_fall(oracle):
    """sAccelerateMove! is exact for constant acceleration: y = y0 + g t^2 / 2."""
    assert free_fall_error(oracle) < 1e-4


def test_free_fall_catches_full_step_velocity(oracle):
    assert free_fall_error(oracle, oracle.mutant(7)) > 1e-2


def test_tait_eos_and_volume(oracle):
    s = pair_system(oracle)
    s.FP[:2, s.L.Density] = 1000.0
    s.FP[:2, s.L.dDensity] = [0.0, 1e5]
    ph = phys(oracle, dt=1e-4)
    s.selfaction(1, ph)
    rho = s.FP[:2, s.L.Density].astype(np.float64)
    assert rho[0] == 1000.0 and rho[1] == pytest.approx(1010.0, rel=1e-6)
    p = s.FP[:2, s.L.Pressure]
    assert p[0] == 0.0                                              # EOS maps rho0 -> p0
    assert p[1] == pytest.approx(ph.tait_b * (1.01 ** 7 - 1), rel=1e-5)
    assert np.allclose(s.FP[:2, s.L.Volume], s.FP[:2, s.L.Mass] / s.FP[:2, s.L.Density], rtol=1e-7)
    assert np.all(s.FP[:2, s.L.dDensity] == 0)


# ------------------------------------------------------------------ search

def brute_force_pairs(x, y, h, tag, adj):
    """O(N^2) symmetry criterion with the reference's FP32 operation order."""
    n = x.shape[0]
    out = []
    for i in range(n):
        dx = (x[i] - x).astype(np.float32)
        dy = (y[i] - y).astype(np.float32)
        dr2 = (dx * dx + dy * dy).astype(np.float32)
        hm = (np.float32(2.0) * h[i] * h / (h[i] + h)).astype(np.float32)
        h2 = (hm * np.float32(2.0)).astype(np.float32)
        ok = (dr2 <= h2 * h2) & (np.arange(n) != i) & (adj[tag[i] - 1, tag - 1] == 1)
        out.append(np.nonzero(ok)[0])
    return out


def cell_list_matches_brute_force(O, variant=None, jitter=0.0):
    nx, ny = 24, 20
    s, x, y = lattice(O, nx, ny, 0.01, h_factor=2.0, jitter=jitter, variant=variant)
    s.search(O.CRIT_SYMMETRY, 0)
    h = s.FP[: s.n, s.L.H]
    tag = s.IP[: s.n, s.L.Tag]
    bf = brute_force_pairs(x, y, h, tag, np.array([[1, 1], [1, 0]]))
    return all(sorted((s.neighbours(i) - 1).tolist()) == bf[i].tolist() for i in range(s.n))


@pytest.mark.parametrize("jitter", [0.0, 0.3])
def test_cell_list_equals_brute_force(oracle, jitter):
    """On a lattice 4 of 48 neighbours sit exactly on r = 2h; jittered lattices
    cover the generic case."""
    assert cell_list_matches_brute_force(oracle, jitter=jitter)


def test_brute_force_catches_strict_criterion(oracle):
    assert not cell_list_matches_brute_force(oracle, oracle.mutant(3))


def test_search_is_permutation_invariant(oracle):
    s, x, y = lattice(oracle, 20, 20, 0.01, jitter=0.4, seed=3)
    s.search(oracle.CRIT_SYMMETRY, 0)
    ref = [set((s.neighbours(i) - 1).tolist()) for i in range(s.n)]
    perm = np.random.default_rng(5).permutation(s.n)
    s2 = oracle.System(s.IP[perm], s.FP[perm], s.is_alive[perm], np.zeros(s.n, np.int32), s.n, s.L, s.domain,
                       np.array([[1, 1], [1, 0]]), M)
    s2.search(oracle.CRIT_SYMMETRY, 0)
    for k in range(s.n):
        got = {int(perm[j]) for j in (s2.neighbours(k) - 1)}
        assert got == ref[perm[k]]


# ------------------------------------------------------------- dam break

def test_small_dam_break_case_and_conservation(oracle):
    """dx = 0.02 case: 6,518 particles, 153 outside the domain (the lattice
    quirk of Geometry2D.jl:83), total mass constant over steps."""
    c = oracle.dam_break_case(0.02, 1.5, 40)
    assert c.n_particles == 6518
    s = oracle.System(c.IP, c.FP, c.is_alive, c.cell_index, c.n_particles, c.layout, oracle.domain2d(*c.domain_args),
                      np.array([[1, 1], [1, 0]]), 40)
    c0 = 10 * math.sqrt(2 * 9.81 * 2.0)
    ph = phys(oracle, kernel=0, dt=np.float32(0.1 * 0.03 / c0), c0=c0, tait_b=1000 * c0 * c0 / 7, av_c=c0)
    assert s.search(oracle.CRIT_SYMMETRY, 1, ph) == 0
    assert int((s.is_alive == 0).sum()) == 153
    mass0 = s.FP[: s.n, s.L.Mass][s.is_alive == 1].astype(np.float64).sum()
    err, pairs = s.run_steps(50, ph)
    assert err == 0 and int((s.is_alive == 0).sum()) == 153
    assert s.FP[: s.n, s.L.Mass][s.is_alive == 1].astype(np.float64).sum() == mass0
    fluid = (s.IP[: s.n, s.L.Tag] == 1) & (s.is_alive == 1)
    y = s.FP[: s.n, s.L.PositionVec + 1][fluid]
    assert y.min() > -0.02 and np.all(np.isfinite(s.FP[: s.n, s.L.Density]))
    walls = s.IP[: s.n, s.L.Tag] == 2
    assert np.array_equal(s.FP[: s.n, s.L.PositionVec][walls], c.FP[: c.n_particles, c.layout.PositionVec][walls])


# ------------------------------------------------------------ Shepard filter
# iKernelFilter! / sKernelFilter! (Filter.jl:10-38): rho_i = (w0 m_i + sum_j W_ij m_j) /
# (w0 V_i + sum_j W_ij V_j) over fluid neighbours j, then V = m / rho and p = eos(rho).

def shepard(O, s, ph):
    assert s.search(O.CRIT_SYMMETRY, 1, ph) == 0
    s.interaction(2, ph)
    s.selfaction(3, ph)
    return s.FP[: s.n, s.L.Density].astype(np.float64)


def wall_row_lattice(O, variant=None):
    """Fluid lattice whose bottom three rows are walls five times denser."""
    nx, ny, dx = 20, 20, 0.01
    tags = np.ones(nx * ny, np.int32)
    tags[: 3 * nx] = 2
    s, x, y = lattice(O, nx, ny, dx, tags=tags, variant=variant)
    wall = tags == 2
    s.FP[: s.n, s.L.Density][wall] *= 5.0
    s.FP[: s.n, s.L.Mass][wall] *= 5.0
    return s, ~wall


def test_shepard_filter_uniform_and_walls_excluded(oracle):
    """A uniform fluid stays at rho0 (up to rounding), also next to denser walls."""
    s, fluid = wall_row_lattice(oracle)
    rho = shepard(oracle, s, phys(oracle))
    assert np.max(np.abs(rho[fluid] / 1000.0 - 1.0)) < 1e-6
    assert s.FP[: s.n, s.L.SumWeight].max() == 0.0  # sums are reset


def test_shepard_filter_catches_wall_neighbours(oracle):
    s, fluid = wall_row_lattice(oracle, variant=oracle.mutant(8))
    rho = shepard(oracle, s, phys(oracle))
    assert np.max(np.abs(rho[fluid] / 1000.0 - 1.0)) > 1e-2


def two_particle_shepard(O, variant=None):
    """Two fluid particles at r = 0.7 h with masses rho0 V and 3 rho0 V: particle 1 becomes
    rho0 (1 + 3 k) / (1 + k), k = W(q) / W(0) = (2 - q)^4 (1 + 2 q) / 16 (Wendland C2, 2D)."""
    s, x, y = lattice(O, 2, 1, 0.01, variant=variant)
    h = float(s.FP[0, s.L.H])
    s.FP[1, s.L.PositionVec] = s.FP[0, s.L.PositionVec] + np.float32(0.7 * h)
    s.FP[1, s.L.Mass] *= 3.0
    s.FP[1, s.L.Density] *= 3.0
    rho = shepard(O, s, phys(O))
    q = float(s.FP[1, s.L.PositionVec] - s.FP[0, s.L.PositionVec]) / h
    k = (2.0 - q) ** 4 * (1.0 + 2.0 * q) / 16.0
    return rho[0], 1000.0 * (1.0 + 3.0 * k) / (1.0 + k)


def test_shepard_filter_two_particles(oracle):
    got, want = two_particle_shepard(oracle)
    assert abs(got / want - 1.0) < 1e-6


def test_shepard_filter_catches_dropped_self_term(oracle):
    got, want = two_particle_shepard(oracle, variant=oracle.mutant(9))
    assert not abs(got / want - 1.0) < 1e-3


def test_shepard_filter_scales_with_mass(oracle):
    """Scaling every mass (and density) by 3 scales the filtered density by 3."""
    a, _ = wall_row_lattice(oracle)
    b, _ = wall_row_lattice(oracle)
    b.FP[: b.n, b.L.Mass] *= 3.0
    b.FP[: b.n, b.L.Density] *= 3.0
    ra, rb = shepard(oracle, a, phys(oracle)), shepard(oracle, b, phys(oracle))
    assert np.max(np.abs(rb / (3.0 * ra) - 1.0)) < 1e-6


# ------------------------------------------------ extrapolated wall pressure

def hydrostatic_walls(O, variant=None, wall_rows=6):
    """A 20 x 20 lattice whose bottom rows are walls; the fluid above holds the
    hydrostatic pressure p = rho0 g (H - y) of a free surface at H."""
    nx, ny, dx = 20, 20, 0.01
    tags = np.ones(nx * ny, np.int32)
    tags[: wall_rows * nx] = 2
    s, x, y = lattice(O, nx, ny, dx, tags=tags, variant=variant)
    H = ny * dx
    fluid = tags == 1
    s.FP[: s.n, s.L.Pressure][fluid] = (1000.0 * 9.81 * (H - y[fluid])).astype(np.float32)
    return s, x, y, fluid, H


def extrapolate(O, s, ph):
    assert s.search(O.CRIT_SYMMETRY, 1, ph) == 0
    s.interaction(3, ph)
    s.selfaction(4, ph)
    return s.FP[: s.n, s.L.Pressure].astype(np.float64)


def test_extrapolated_wall_pressure_is_hydrostatic(oracle):
    """Walls within reach of the fluid take rho0 g (H - y_wall), the hydrostatic
    pressure carried down to their depth; walls out of reach keep p0 and their
    sums; the sums of the others are cleared."""
    s, x, y, fluid, H = hydrostatic_walls(oracle)
    p = extrapolate(oracle, s, phys(oracle))
    wall = ~fluid
    # fluid starts at y = 6.5 dx and W vanishes at 2h = 4 dx, so rows from 3.5 dx up see it
    reach = wall & (y > 0.03)
    want = 1000.0 * 9.81 * (H - y)
    assert reach.sum() > 0 and (wall & ~reach).sum() > 0
    assert np.max(np.abs(p[reach] / want[reach] - 1.0)) < 1e-5
    assert (p[wall & ~reach] == 0.0).all()
    assert (s.FP[: s.n, s.L.SumWeight] == 0.0).all() and (s.FP[: s.n, s.L.SumWeightedPressure] == 0.0).all()


def test_extrapolated_wall_pressure_catches_flipped_gravity(oracle):
    s, x, y, fluid, H = hydrostatic_walls(oracle, variant=oracle.mutant(10))
    p = extrapolate(oracle, s, phys(oracle))
    reach = ~fluid & (y > 0.03)
    want = 1000.0 * 9.81 * (H - y)
    assert np.max(np.abs(p[reach] / want[reach] - 1.0)) > 1e-2


def test_extrapolated_wall_pressure_catches_uncleared_sums(oracle):
    s, *_ = hydrostatic_walls(oracle, variant=oracle.mutant(11))
    extrapolate(oracle, s, phys(oracle))
    assert s.FP[: s.n, s.L.SumWeight].max() > 0.0


def test_extrapolated_wall_pressure_clamps_and_propagates_nan(oracle):
    """max(p_j, p0): suction in the fluid gives p0 on the wall; Julia's max
    keeps a NaN, so a NaN fluid pressure reaches the walls next to it only."""
    s, x, y, fluid, H = hydrostatic_walls(oracle)
    s.FP[: s.n, s.L.Pressure][fluid] = -500.0
    p = extrapolate(oracle, s, phys(oracle, gy=0.0))
    assert (p[~fluid] == 0.0).all()
    s, x, y, fluid, H = hydrostatic_walls(oracle)
    k = int(np.flatnonzero(fluid & (y < 0.07) & (x > 0.09) & (x < 0.10))[0])  # first fluid row, mid x
    s.FP[k, s.L.Pressure] = np.nan
    p = extrapolate(oracle, s, phys(oracle))
    near = ~fluid & ((x - x[k]) ** 2 + (y - y[k]) ** 2 < (0.04) ** 2)
    far = ~fluid & (np.abs(x - x[k]) > 0.05)
    assert np.isnan(p[near]).any() and np.isfinite(p[far]).all()


# ------------------------------------- the reference example's own physics
# main.jl:57-70: linear EOS p = c0^2 (rho - rho0) + p0 with c0 = 120, and
# iClassicViscosity! with mu0 (main.jl:110,114 -> Viscosity.jl:10-26).

def linear_eos_values(O, variant=None):
    s = pair_system(O, variant)
    s.FP[:2, s.L.Density] = 1000.0
    s.FP[:2, s.L.dDensity] = [0.0, 1.0e4]                # rho -> 1000 and 1001 with dt = 1e-4
    ph = phys(O, eos=0, c0=120.0, p0=250.0, dt=1e-4)
    s.selfaction(1, ph)
    return s.FP[:2, s.L.Density].copy(), s.FP[:2, s.L.Pressure].copy()


def test_linear_eos_closed_form(oracle):
    """rho0 maps to p0 exactly; one unit of density above rho0 adds c0^2."""
    rho, p = linear_eos_values(oracle)
    assert rho.tolist() == [1000.0, 1001.0]
    assert p[0] == 250.0
    assert p[1] == 250.0 + 120.0 * 120.0


def test_linear_eos_check_catches_unsquared_sound_speed(oracle):
    _, p = linear_eos_values(oracle, oracle.mutant(12))
    assert p[1] != 250.0 + 120.0 * 120.0


def classic_viscosity_residual(O, variant=None):
    """p = 0: classic viscosity alone; the pair's exchange m_i du_i + m_j du_j cancels."""
    s = pair_system(O, variant, p=(0.0, 0.0))
    ph = phys(O, visc=0, mu=1e-3)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    s.interaction(1, ph)
    du = s.FP[:2, s.L.dVelocityVec: s.L.dVelocityVec + 2].astype(np.float64)
    m = s.FP[:2, s.L.Mass].astype(np.float64)
    total = m[0] * du[0] + m[1] * du[1]
    return np.max(np.abs(total)) / np.max(np.abs(m[:, None] * du)), du


def test_classic_viscosity_antisymmetric(oracle):
    res, du = classic_viscosity_residual(oracle)
    assert np.max(np.abs(du)) > 0.0 and res < 1e-5
    u_rel = np.array([0.3, 0.1]) - np.array([-0.2, 0.05])
    assert float(np.dot(du[0], u_rel)) < 0.0                     # it slows the relative motion


def test_classic_viscosity_antisymmetry_catches_transposed_mass(oracle):
    assert classic_viscosity_residual(oracle, oracle.mutant(14))[0] > 1e-2


def shear_decay_ratio(O, variant=None):
    """u_x = U sin(k y) on a lattice, p = 0, rho = rho0: the SPH classic-viscosity
    acceleration is nu times the Laplacian (Morris' form), so interior particles
    decay at du_x / u_x = -nu k^2 with nu = mu / rho0.  Returns the median of
    (du_x / u_x) / (-nu k^2) over interior particles away from the nodes."""
    nx, ny = 40, 80
    dx, lam, U, mu = 0.01, 40 * 0.01, 0.5, 1e-3
    k = 2 * math.pi / lam
    s, x, y = lattice(O, nx, ny, dx, h_factor=1.5, variant=variant,
                      vel=lambda x, y: ((U * np.sin(k * y)).astype(np.float32), 0 * y))
    ph = phys(O, kernel=0, visc=0, mu=mu, eos=0, c0=120.0)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    s.interaction(1, ph)
    du = s.FP[: s.n, s.L.dVelocityVec].astype(np.float64)
    dv = s.FP[: s.n, s.L.dVelocityVec + 1].astype(np.float64)
    u = s.FP[: s.n, s.L.VelocityVec].astype(np.float64)
    sel = interior(x, y, nx, ny, dx) & (np.abs(u) > 0.5 * U)
    assert sel.sum() > 100
    assert np.max(np.abs(dv[sel])) < 1e-3 * np.max(np.abs(du[sel]))   # no cross-stream acceleration
    return float(np.median((du[sel] / u[sel]) / (-(mu / 1000.0) * k * k)))


def test_classic_viscosity_shear_layer_decay_rate(oracle):
    assert abs(shear_decay_ratio(oracle) - 1.0) < 0.05


def test_shear_decay_catches_dropped_factor_two(oracle):
    assert abs(shear_decay_ratio(oracle, oracle.mutant(13)) - 1.0) > 0.3


# ------------------------------------- sum-of-|terms| and summation order
# (the FAST-mode bars of the GPU tests are 1e-5 of these magnitudes)

def wendland_sensitivity(q):
    """(2 |G| + |q G'|) / |G| for the Wendland C2 gradient shape G = q (2 - q)^3."""
    g = q * (2 - q) ** 3
    gp = (2 - q) ** 2 * (2 - 4 * q)
    return (2 * abs(g) + abs(q * gp)) / abs(g)


def test_term_magnitude_of_a_single_pair_is_its_term(oracle):
    """One neighbour each, motion along x (no inner cancellation), p = 0: du
    is one artificial-viscosity term and dDensity one continuity term, both
    proportional to f = dw / r, so their bounds are |term| (|f| + r |f'|) / |f|
    = |term| (2 |G| + |q G'|) / |G| (closed form of the Wendland shape; with
    a flat EOS, so that the pressure share of one density ulp is zero); with
    pressure the bound is at least |du|."""
    O = oracle
    ph = phys(O)
    s = pair_system(O, p=(0.0, 0.0), u=((0.3, 0.0), (-0.3, 0.0)))
    s.search(O.CRIT_SYMMETRY, 1, ph)
    q = float(s.FP[0, s.L.nR]) * float(s.FP[0, s.L.nHInv])
    k = wendland_sensitivity(q)
    c = s.term_magnitude(0, ph)
    drho = np.abs(s.FP[:2, s.L.dDensity].astype(np.float64))
    assert drho.min() > 0 and np.allclose(c / drho, k, rtol=1e-5)
    flat = phys(O, tait_b=0.0)  # dp / drho = 0: no pressure-ulp share in the bound
    a = s.term_magnitude(1, flat)
    s.interaction(1, flat)
    du = np.abs(s.FP[:2, s.L.dVelocityVec: s.L.dVelocityVec + 2].astype(np.float64))
    assert du[:, 0].min() > 0 and np.allclose(a[:, 0] / du[:, 0], k, rtol=1e-5)
    s = pair_system(O)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    a = s.term_magnitude(1, ph)
    s.interaction(1, ph)
    du = np.abs(s.FP[:2, s.L.dVelocityVec: s.L.dVelocityVec + 2].astype(np.float64))
    assert (a >= du).all() and a.max() > 0


def test_term_magnitude_grows_at_the_support_edge(oracle):
    """Near q = 2 the Wendland gradient's relative sensitivity to r grows like
    3 / (2 - q): a pair at q = 1.999 gets a bound ~1000x its term."""
    O = oracle
    ph = phys(O)
    s = pair_system(O, p=(0.0, 0.0), u=((0.3, 0.0), (-0.3, 0.0)))
    hinv = float(O.lib().or_invharmonic(0.010, 0.011))
    r = 1.999 / hinv
    s.FP[1, s.L.PositionVec] = np.float32(0.100 + r)
    s.FP[1, s.L.PositionVec + 1] = np.float32(0.100)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    assert s.count().tolist() == [1, 1]
    q = float(s.FP[0, s.L.nR]) * float(s.FP[0, s.L.nHInv])
    c = s.term_magnitude(0, ph)
    drho = np.abs(s.FP[:2, s.L.dDensity].astype(np.float64))
    assert wendland_sensitivity(q) > 500 and np.allclose(c / drho, wendland_sensitivity(q), rtol=1e-3)


def test_reversed_order_changes_only_rounding(oracle):
    """A lattice with a velocity field and a hydrostatic pressure: momentum
    summed in reversed neighbour order differs from the list order by
    rounding only (within 1e-6 of sum |terms|), and not at all for the
    two-particle system."""
    O = oracle
    nx = ny = 24
    dx = 0.01
    s, x, y = lattice(O, nx, ny, dx, vel=lambda x, y: ((0.3 * np.sin(9 * y)).astype(np.float32),
                                                       (0.2 * np.cos(7 * x)).astype(np.float32)),
                      p=lambda x, y: (1e4 * (1.0 - y)).astype(np.float32))
    ph = phys(O)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    a = s.term_magnitude(1, ph)
    c = slice(s.L.dVelocityVec, s.L.dVelocityVec + 2)
    before = s.FP[:, c].copy()
    s.interaction(1, ph)
    fwd = s.FP[: s.n, c].astype(np.float64)
    s.FP[:, c] = before
    s.set_pair_order(1)
    try:
        s.interaction(1, ph)
    finally:
        s.set_pair_order(0)
    rev = s.FP[: s.n, c].astype(np.float64)
    assert (fwd != rev).any()                                      # the order matters to the bits
    assert np.max(np.abs(fwd - rev) / a) < 1e-6
    p2 = pair_system(O)
    p2.search(O.CRIT_SYMMETRY, 1, ph)
    p2.set_pair_order(1)
    try:
        p2.interaction(1, ph)
    finally:
        p2.set_pair_order(0)
    rev2 = p2.FP[:2, c].copy()
    p3 = pair_system(O)
    p3.search(O.CRIT_SYMMETRY, 1, ph)
    p3.interaction(1, ph)
    assert np.array_equal(rev2, p3.FP[:2, c])
