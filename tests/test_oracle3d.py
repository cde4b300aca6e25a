"""Pins the 3D restatement (oracle/eppx_oracle3d.c).  The reference's 3D
layers are stubs (NeighbourSystemBase.jl:50-57, NeighbourSearch.jl:229,
AbstractDomain.jl:119), so nothing from the reference pins it: these are
closed forms, invariants and brute force, each paired with a restatement
variant that plants one plausible 3D mistake (OR_MUTATION 15-21) and must
make its check fail."""
import math

import numpy as np
import pytest

M = 300


@pytest.fixture(scope="module")
def o3(oracle):
    import oracle3d
    return oracle3d


def layout3(O, n, M=M):
    int_nt = {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": M}
    float_nt = {"PositionVec": 3, "VelocityVec": 3, "dVelocityVec": 3, "Mass": 1, "Volume": 1, "Density": 1,
                "dDensity": 1, "Pressure": 1, "Gap": 1, "H": 1, "nW": M, "nDW": M, "nHInv": M, "nRVec": 3 * M,
                "nR": M}
    return O.make_layout(int_nt, float_nt, n)


def lattice3(O, o3, nx, ny, nz, dx, h_factor=2.0, jitter=0.0, seed=0, tags=None, vel=None, rho=None, p=None,
             variant=None, h_jitter=0.0):
    """nx x ny x nz lattice of fluid particles, cell size 2h, margin 2h."""
    n = nx * ny * nz
    L = layout3(O, n)
    IP = np.zeros((n, L.n_int), np.int32, order="F")
    FP = np.zeros((n, L.n_float), np.float32, order="F")
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([(0.5 + i.ravel()) * dx, (0.5 + j.ravel()) * dx, (0.5 + k.ravel()) * dx], axis=1)
    rng = np.random.default_rng(seed)
    if jitter:
        pos = pos + rng.uniform(-jitter, jitter, pos.shape) * dx
    pos = pos.astype(np.float32)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    for c in range(3):
        FP[:, L.PositionVec + c] = pos[:, c]
    IP[:, L.Tag] = 1 if tags is None else tags
    IP[:, L.IsMovable] = 1
    vol = np.float32(dx * dx * dx)
    FP[:, L.Volume] = vol
    FP[:, L.Density] = 1000.0 if rho is None else rho(x, y, z)
    FP[:, L.Mass] = FP[:, L.Density] * vol
    h = np.float32(h_factor * dx)
    FP[:, L.H] = h if not h_jitter else (h * (1 + rng.uniform(-h_jitter, h_jitter, n))).astype(np.float32)
    if vel is not None:
        for c, comp in enumerate(vel(x, y, z)):
            FP[:, L.VelocityVec + c] = comp
    if p is not None:
        FP[:, L.Pressure] = p(x, y, z)
    gap = 2 * h_factor * dx * (1 + h_jitter)
    d = o3.domain3d(gap, -gap, -gap, -gap, nx * dx + gap, ny * dx + gap, nz * dx + gap)
    adj = np.array([[1, 1], [1, 0]], np.int32)
    s = o3.System3D(IP, FP, np.ones(n, np.int32), np.zeros(n, np.int32), n, L, d, adj, M, variant=variant)
    return s, x, y, z


def phys3(O, o3, kernel=1, gz=-9.81, **kw):
    base = dict(fluid_tag=1, wall_tag=2, kernel=kernel, eos=1, visc=1, dt=1e-4, rho0=1000.0, p0=0.0, c0=44.3,
                tait_b=1000.0 * 44.3 ** 2 / 7, gx=0.0, gy=0.0, alpha=0.1, beta=0.1, av_c=44.3, mu=1e-3,
                coefficient=0.0)
    base.update(kw)
    return o3.physics3d(gz=gz, **base)


def interior3(x, y, z, nx, ny, nz, dx, margin=4):
    return ((x > margin * dx) & (x < (nx - margin) * dx) & (y > margin * dx) & (y < (ny - margin) * dx)
            & (z > margin * dx) & (z < (nz - margin) * dx))


# ------------------------------------------------------------------ kernels

def normalisation(lib, kernel, h=0.01):
    r = np.linspace(0, 2 * h, 20001)
    w = np.array([lib.or3_kernel_value(kernel, float(v), 1.0 / h) for v in r])
    return np.trapezoid(4 * math.pi * r * r * w, r)


@pytest.mark.parametrize("kernel", [0, 1])
def test_kernel_3d_normalisation_support_and_centre(o3, kernel):
    lib = o3.lib()
    h = 0.01
    assert normalisation(lib, kernel, h) == pytest.approx(1.0, rel=2e-4)      # ∫ W dV = 1
    assert lib.or3_kernel_value(kernel, 2 * h, 1.0 / h) == 0.0               # compact support
    assert lib.or3_kernel_gradient(kernel, 2 * h * 1.0001, 1.0 / h) == 0.0
    sigma = 1 / math.pi if kernel == 0 else 21 / 16 / math.pi                # CubicSpline.jl:15, WendlandC2.jl:15
    centre = sigma / h ** 3 * (1.0 if kernel == 0 else 1.0)
    assert lib.or3_kernel_value(kernel, 0.0, 1.0 / h) == pytest.approx(centre, rel=1e-6)
    assert lib.or3_kernel_value0(kernel, h) == pytest.approx(sigma / h ** 3, rel=1e-6)


def test_normalisation_catches_2d_sigma(o3):
    assert abs(normalisation(o3.mutant(16), 1) - 1.0) > 0.1


def gradient3_matches_derivative(lib, kernel):
    h = 0.01
    for q in (0.2, 0.7, 0.99, 1.01, 1.3, 1.8):
        r, e = q * h, 1e-3 * h  # float32 values: a wider step keeps the rounding of W out of the quotient
        fd = (lib.or3_kernel_value(kernel, r + e, 1 / h) - lib.or3_kernel_value(kernel, r - e, 1 / h)) / (2 * e)
        if abs(lib.or3_kernel_gradient(kernel, r, 1 / h) - fd) > 2e-3 * abs(fd):
            return False
    return True


@pytest.mark.parametrize("kernel", [0, 1])
def test_kernel_3d_gradient_is_derivative(o3, kernel):
    assert gradient3_matches_derivative(o3.lib(), kernel)


def test_gradient_check_catches_hinv_cubed(o3):
    assert not gradient3_matches_derivative(o3.mutant(17), 1)


# ---------------------------------------------------------- domain, stencil

def test_domain3d_cells_and_locate(o3):
    d = o3.domain3d(0.1, -0.1, -0.2, -0.3, 1.0, 0.5, 0.45)
    assert (d.n_x, d.n_y, d.n_z) == (11, 7, 7) and d.n == 11 * 7 * 7
    lib = o3.lib()
    # cell (i, j, k) of a point, 1-based, x fastest
    x, y, z = np.float32(0.35), np.float32(-0.05), np.float32(0.4)
    i = math.floor((x - d.first_x) * d.gap_x_inv) + 1
    j = math.floor((y - d.first_y) * d.gap_y_inv) + 1
    k = math.floor((z - d.first_z) * d.gap_z_inv) + 1
    assert lib.or3_index_linear_from_position(d, x, y, z) == i + d.n_x * ((j - 1) + d.n_y * (k - 1))
    assert lib.or3_inside(d, x, y, z) == 1 and lib.or3_inside(d, x, y, np.float32(0.46)) == 0
    # positions on the upper faces clamp into the last cell
    assert lib.or3_index_linear_from_position(d, d.last_x, d.last_y, d.last_z) == d.n


def stencil_ok(o3, variant=None):
    d = o3.domain3d(1.0, 0.0, 0.0, 0.0, 4.0, 3.0, 5.0)
    s = o3.System3D(np.zeros((1, 1), np.int32), np.zeros((1, 1), np.float32), np.zeros(1, np.int32),
                    np.zeros(1, np.int32), 0, layout3(__import__("oracle"), 1), d, np.ones((1, 1), np.int32), 4,
                    variant=variant)
    nc = d.n
    cnt = np.ctypeslib.as_array(s.ns.neighbour_cell_count, shape=(nc,))
    lst = np.ctypeslib.as_array(s.ns.neighbour_cell_list, shape=(27, nc)).T
    for I in range(1, nc + 1):
        c = I - 1
        i, j, k = c % d.n_x + 1, (c // d.n_x) % d.n_y + 1, c // (d.n_x * d.n_y) + 1
        want = [ii + d.n_x * ((jj - 1) + d.n_y * (kk - 1))
                for ii in (i - 1, i, i + 1) for jj in (j - 1, j, j + 1) for kk in (k - 1, k, k + 1)
                if 1 <= ii <= d.n_x and 1 <= jj <= d.n_y and 1 <= kk <= d.n_z]
        if cnt[c] != len(want) or lst[c, : cnt[c]].tolist() != want:
            return False
    return True


def test_stencil_order_and_counts(o3):
    """27 cells inside, 8 at a corner; x offset outer, then y, then z (NeighbourSystemBase.jl:37-47 with z)."""
    assert stencil_ok(o3)


def test_stencil_check_catches_missing_layer(o3):
    assert not stencil_ok(o3, o3.mutant(18))


# ------------------------------------------------------------------ search

def brute_force_pairs3(pos, h, tag, adj):
    """O(N^2) symmetry criterion, FP32, dr2 = dx*dx + dy*dy + dz*dz left to right."""
    n = pos.shape[0]
    out = []
    for i in range(n):
        d = (pos[i] - pos).astype(np.float32)
        dr2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)
        hm = (np.float32(2.0) * h[i] * h / (h[i] + h)).astype(np.float32)
        h2 = (hm * np.float32(2.0)).astype(np.float32)
        ok = (dr2 <= h2 * h2) & (np.arange(n) != i) & (adj[tag[i] - 1, tag - 1] == 1)
        out.append(np.nonzero(ok)[0])
    return out


def cell_list_matches_brute_force3(O, o3, variant=None, jitter=0.0, h_jitter=0.0, seed=0):
    nx, ny, nz = 11, 9, 10
    tags = np.where(np.random.default_rng(seed + 7).uniform(size=nx * ny * nz) < 0.3, 2, 1).astype(np.int32)
    s, x, y, z = lattice3(O, o3, nx, ny, nz, 0.01, jitter=jitter, seed=seed, tags=tags, variant=variant,
                          h_jitter=h_jitter)
    if s.search(O.CRIT_SYMMETRY, 0) != 0:  # a neighbour list overflowed: not the brute-force set
        return False
    pos = np.stack([x, y, z], axis=1)
    h = s.FP[: s.n, s.L.H]
    bf = brute_force_pairs3(pos, h, tags, np.array([[1, 1], [1, 0]]))
    return all(sorted((s.neighbours(i) - 1).tolist()) == bf[i].tolist() for i in range(s.n))


@pytest.mark.parametrize("jitter,h_jitter", [(0.0, 0.0), (0.3, 0.0), (0.3, 0.03)])
def test_cell_list_equals_brute_force_3d(oracle, o3, jitter, h_jitter):
    """Lattice (6 exact ties at r = 2h per interior particle), jittered, and jittered with varying H; fluid and
    wall tags (wall-wall pairs are not adjacent)."""
    assert cell_list_matches_brute_force3(oracle, o3, jitter=jitter, h_jitter=h_jitter)


@pytest.mark.parametrize("mutation", [15, 18])
def test_brute_force_catches_dropped_z(oracle, o3, mutation):
    assert not cell_list_matches_brute_force3(oracle, o3, o3.mutant(mutation), jitter=0.3)


def test_lattice_neighbour_count(oracle, o3):
    """On a cubic lattice with h = 2dx an interior particle has the 250 lattice points of |r| < 4dx as
    neighbours, plus those of the 6 points on the support radius |r| = 4dx that the float32 rounding of the
    positions puts inside (the 3D count of 256 for r <= 4dx)."""
    s, x, y, z = lattice3(oracle, o3, 14, 14, 14, 0.01)
    s.search(oracle.CRIT_SYMMETRY, 0)
    sel = interior3(x, y, z, 14, 14, 14, 0.01, margin=5)
    cnt = s.count()[sel]
    assert sel.sum() > 0 and cnt.min() >= 250 and cnt.max() <= 256
    pts = np.array([(a, b, c) for a in range(-4, 5) for b in range(-4, 5) for c in range(-4, 5)])
    r2 = (pts ** 2).sum(axis=1)
    assert ((r2 > 0) & (r2 < 16)).sum() == 250 and (r2 == 16).sum() == 6


# -------------------------------------------------------------- continuity

def divergence3_error(O, o3, variant=None, axis=2):
    """u_axis = a x_axis: interior drho ≈ -rho0 a (dρ/dt = -ρ ∇·u)."""
    nx = ny = nz = 16
    dx, a = 0.01, 0.5
    vel = lambda x, y, z: tuple((a * (x, y, z)[c] if c == axis else 0 * x) for c in range(3))
    s, x, y, z = lattice3(O, o3, nx, ny, nz, dx, vel=vel, variant=variant)
    assert s.search(O.CRIT_SYMMETRY, 1, phys3(O, o3)) == 0
    drho = s.FP[: s.n, s.L.dDensity]
    sel = interior3(x, y, z, nx, ny, nz, dx)
    return np.max(np.abs(drho[sel] - (-1000.0 * a)) / (1000.0 * a))


@pytest.mark.parametrize("axis", [0, 2])
def test_continuity_3d_divergence(oracle, o3, axis):
    assert divergence3_error(oracle, o3, axis=axis) < 0.03


def test_divergence_catches_dropped_z_velocity(oracle, o3):
    assert divergence3_error(oracle, o3, o3.mutant(21), axis=2) > 0.5


# ---------------------------------------------------------------- momentum

def test_momentum_3d_antisymmetric(oracle, o3):
    """Pressure + artificial viscosity between fluid particles: Σ m_i du_i = 0 over a cluster with random
    velocities, densities and pressures."""
    rng = np.random.default_rng(4)
    n = 10 * 9 * 8
    rho = lambda x, y, z: (1000.0 + rng.uniform(-20, 20, x.shape)).astype(np.float32)
    p = lambda x, y, z: rng.uniform(0, 3e4, x.shape).astype(np.float32)
    vel = lambda x, y, z: tuple(rng.uniform(-0.5, 0.5, x.shape).astype(np.float32) for _ in range(3))
    s, x, y, z = lattice3(oracle, o3, 10, 9, 8, 0.01, jitter=0.2, rho=rho, p=p, vel=vel)
    ph = phys3(oracle, o3)
    s.search(oracle.CRIT_SYMMETRY, 1, ph)
    s.interaction(1, ph)
    du = s.FP[:n, s.L.dVelocityVec: s.L.dVelocityVec + 3].astype(np.float64)
    m = s.FP[:n, s.L.Mass].astype(np.float64)
    total = (m[:, None] * du).sum(axis=0)
    assert np.max(np.abs(du)) > 1.0
    assert np.max(np.abs(total)) < 1e-5 * np.abs(m[:, None] * du).sum(axis=0).max()


def hydrostatic_error(O, o3, variant=None):
    """p = rho0 g (H - z), rho = rho0: interior du ≈ (0, 0, +g) (the pressure gradient that balances gravity)."""
    nx = ny = nz = 16
    dx, g = 0.01, 9.81
    H = nz * dx
    s, x, y, z = lattice3(O, o3, nx, ny, nz, dx, p=lambda x, y, z: (1000.0 * g * (H - z)).astype(np.float32),
                          variant=variant)
    ph = phys3(O, o3)
    s.search(O.CRIT_SYMMETRY, 1, ph)
    s.interaction(1, ph)
    du = s.FP[: s.n, s.L.dVelocityVec: s.L.dVelocityVec + 3]
    sel = interior3(x, y, z, nx, ny, nz, dx)
    return max(float(np.max(np.abs(du[sel, 2] - g))) / g, float(np.max(np.abs(du[sel, :2]))) / g)


def test_momentum_3d_hydrostatic(oracle, o3):
    assert hydrostatic_error(oracle, o3) < 0.03


def test_hydrostatic_catches_dropped_z_pressure_term(oracle, o3):
    assert hydrostatic_error(oracle, o3, o3.mutant(20)) > 0.5


# -------------------------------------------------------------- self actions

def free_fall3_error(O, o3, variant=None, steps=20):
    s, x, y, z = lattice3(O, o3, 1, 1, 2, 0.5, variant=variant)  # two particles out of each other's reach
    dt = np.float32(1e-3)
    ph = phys3(O, o3, dt=dt, gz=-9.81)
    z0 = s.FP[:2, s.L.PositionVec + 2].astype(np.float64)
    y0 = s.FP[:2, s.L.PositionVec + 1].astype(np.float64)
    for _ in range(steps):
        s.selfaction(2, ph)
    t = steps * float(dt)
    fall = 0.5 * 9.81 * t * t
    ez = np.abs(s.FP[:2, s.L.PositionVec + 2] - (z0 - fall)) / fall
    ey = np.abs(s.FP[:2, s.L.PositionVec + 1] - y0) / fall
    return float(max(ez.max(), ey.max()))


def test_integrator_3d_free_fall(oracle, o3):
    assert free_fall3_error(oracle, o3) < 1e-4


def test_free_fall_catches_swapped_gravity_axes(oracle, o3):
    assert free_fall3_error(oracle, o3, o3.mutant(19)) > 0.5


def test_small_3d_dam_break_conserves_mass(oracle, o3):
    """20 steps of a small 3D dam break (dx = 0.05): no particle leaves, total mass is constant (Mass is never
    written), every state finite, the block starts to fall (min fluid z velocity < 0)."""
    import sys
    from eppx import cases
    c = cases.dam_break_3d(0.05, M=M, block=(0.5, 0.25, 0.5), tank=(1.2, 0.5, 0.8))
    h = c.host
    L = oracle.make_layout(c.int_nt, c.float_nt, h.n_capacity)
    s = o3.System3D(h.int_properties_, h.float_properties_, h.is_alive_, h.cell_index_, h.n_particles_, L,
                    o3.domain3d(*c.domain.args), np.array([[1, 1], [1, 0]]), M)
    p = c.physics
    ph = o3.physics3d(gz=p.gz, **{k: getattr(p, k) for k in ("fluid_tag", "wall_tag", "kernel", "eos", "visc", "dt",
                                                              "rho0", "p0", "c0", "tait_b", "gx", "gy", "mu",
                                                              "alpha", "beta", "av_c", "coefficient")})
    assert s.search(oracle.CRIT_SYMMETRY, 1, ph) == 0
    m0 = s.FP[: s.n, L.Mass].astype(np.float64).sum()
    err, pairs = s.run_steps(20, ph, threads=4)
    assert err == 0 and np.all(s.is_alive[: s.n] == 1)
    assert s.FP[: s.n, L.Mass].astype(np.float64).sum() == m0
    assert np.all(np.isfinite(s.FP[: s.n, [L.Density, L.Pressure, L.PositionVec + 2, L.VelocityVec + 2]]))
    fluid = s.IP[: s.n, L.Tag] == 1
    assert s.FP[: s.n, L.VelocityVec + 2][fluid].min() < 0.0
    assert pairs.min() > 0 and sys is not None
