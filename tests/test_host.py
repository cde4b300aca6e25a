"""Host-side logic (no GPU): the C-ABI library loads and exports every entry
point include/eppx.h declares, its host arithmetic (Domain2D) matches the
restatement bit for bit, and the Python mirror of the reference's Class /
Geometry / case generation matches the reference's fixtures and the
restatement."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO


@pytest.fixture(scope="module")
def eppx_mod():
    import eppx
    if not os.path.exists(eppx.library_path()):
        eppx.build()
    return eppx


def declared_functions():
    text = open(os.path.join(REPO, "include", "eppx.h")).read()
    return sorted(set(re.findall(r"\b(eppx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(eppx_mod):
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(eppx_mod.library_path())
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert eppx_mod.lib().eppx_abi_version() == 1


DOMAINS = [(0.15, -1.0, -2.0, 3.0, 3.0), (0.15, 0.1, 0.2, 0.9, 0.9), (0.06, -0.06, -0.06, 4.06, 3.06),
           (0.004, -0.004, -0.004, 4.004, 3.004), (0.04, -0.04, -0.04, 4.04, 3.04), (0.01, -0.01, -0.01, 4.01, 3.01)]


@pytest.mark.parametrize("args", DOMAINS)
def test_domain_bits_match_restatement(eppx_mod, oracle, args):
    from eppx import Class
    d = Class.Domain2D(*args)
    r = oracle.domain2d(*args)
    for name, _ in r._fields_:
        a, b = getattr(d.c, name), getattr(r, name)
        assert np.float32(a).tobytes() == np.float32(b).tobytes() if isinstance(a, float) else a == b, name
    rng = np.random.default_rng(0)
    xs = rng.uniform(args[1] - 0.1, args[3] + 0.1, 200).astype(np.float32)
    ys = rng.uniform(args[2] - 0.1, args[4] + 0.1, 200).astype(np.float32)
    for x, y in zip(xs, ys):
        ins = bool(oracle.lib().or_inside(C.byref(r), float(x), float(y)))
        lin = oracle.lib().or_index_linear_from_position(C.byref(r), float(x), float(y))
        assert d.inside(x, y) == ins and d.indexLinearFromPosition(x, y) == lin


def test_domain_golden(eppx_mod, golden):
    from eppx import Class
    g = golden["domain"]
    d = Class.Domain2D(*g["args"])
    assert (d.get_n_x(), d.get_n_y(), d.get_n()) == (g["n_x"], g["n_y"], g["n"])
    assert d.indexCartesianToLinear(*g["cartesian_to_linear"][0]) == g["cartesian_to_linear"][1]
    assert list(d.indexLinearToCartesian(g["linear_to_cartesian"][0])) == g["linear_to_cartesian"][1]
    assert list(d.indexCartesianFromPosition(*g["cartesian_from_position"][0])) == g["cartesian_from_position"][1]
    assert d.indexLinearFromPosition(*g["linear_from_position"][0]) == g["linear_from_position"][1]


def test_named_index_and_basic_index_golden(eppx_mod, golden):
    from eppx import Class
    g = golden["particle_system"]
    ni = Class.NamedIndex(g["int_named_tuple"], g["float_named_tuple"])
    assert (ni.get_n_int_capacity(), ni.get_n_float_capacity()) == (g["n_int_capacity"], g["n_float_capacity"])
    bi = Class.mapBasicIndex(ni, {"PositionVec": "PositionVec", "nCount": "nCount", "nIndex": "nIndex", "nR": "nR"})
    for k, v in g["basic_index"].items():
        assert bi[k] == v, k
    assert len(ni.get_float_symbol_head()) == ni.get_n_float_capacity()
    gn = golden["named_index"]
    assert Class.NamedIndex(gn["int_named_tuple"], g["float_named_tuple"]).get_n_int_capacity() == gn["n_int_capacity"]


def test_active_pair_golden(eppx_mod, golden):
    from eppx import Class
    for g in golden["active_pair"]:
        ap = Class.ActivePair(g["pairs"])
        assert ap.adjacency_matrix_.tolist() == g["adjacency"]


def test_geometry_golden(eppx_mod, golden):
    from eppx import Geometry
    g = golden["geometry"]
    r = Geometry.Rectangle(*g["rectangle"])
    assert Geometry.count(g["gap"], r) == g["count"]
    for (x, y), exp in g["inside"]:
        assert Geometry.inside(x, y, r) == exp


@pytest.mark.parametrize("dx,hf,M", [(0.02, 1.5, 40), (0.01, 2.0, 64), (0.0025, 2.0, 64)])
def test_dam_break_case_matches_restatement(eppx_mod, oracle, dx, hf, M):
    """eppx.cases.dam_break == the restatement of casegen.jl, bit for bit."""
    from eppx import cases
    c = cases.dam_break(dx, hf, M=M)
    r = oracle.dam_break_case(dx, hf, M)
    assert c.n_particles == r.n_particles
    assert np.array_equal(c.host.int_properties_, r.IP)
    assert np.array_equal(c.host.float_properties_.view(np.uint32), r.FP.view(np.uint32))
    assert c.domain.args == r.domain_args


def test_small_case_counts(eppx_mod):
    from eppx import cases
    c = cases.small()
    assert c.n_particles == 6518 and (c.domain.get_n_x(), c.domain.get_n_y()) == (68, 52)
    # the last particle of each lattice row sits at first_x - dx/2 (Geometry2D.jl:83)
    x = c.host.column("PositionVec", 0)[:5000]
    assert np.isclose(x[49], -0.01) and np.isclose(x[0], 0.01)


def test_large_case_counts(eppx_mod):
    from eppx import cases
    c = cases.large()
    assert c.n_particles == 2_030_018
    assert (c.domain.get_n_x(), c.domain.get_n_y()) == (1002, 751)


def test_physics_parameters(eppx_mod):
    from eppx import SPH
    p = SPH.dam_break_physics(0.02, 1.5, SPH.CubicSpline)
    assert p.c0 == pytest.approx(62.6418, rel=1e-5) and p.tait_b == pytest.approx(5.606e5, rel=1e-3)
    assert p.dt == pytest.approx(4.789e-5, rel=1e-3)
    assert SPH.dam_break_physics(0.001, 2.0).dt == pytest.approx(3.193e-6, rel=1e-3)


def test_closures_do_not_cross_the_abi(eppx_mod):
    from eppx import Algorithm
    with pytest.raises(NotImplementedError):
        Algorithm._action(lambda *a: None, "self")
    with pytest.raises(NotImplementedError):
        Algorithm._criterion(lambda *a: True)


def test_mix_and_serving_specs(eppx_mod):
    from eppx import cases
    specs = cases.mix_specs(0)
    assert len(specs) == 64 and {s[0] for s in specs} <= {0.02, 0.01, 0.005, 0.0025}
    assert all(0.5 <= w <= 1.5 and 0.25 <= h <= 3.0 for _, w, h, _ in specs)
    stream = cases.serving_stream(1)
    assert all(1e4 <= n <= 1e6 and 200 <= s <= 2000 for _, n, _, s in stream)
    assert all(b[0] > a[0] for a, b in zip(stream, stream[1:]))
