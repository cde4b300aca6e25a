"""The oracle (CPU restatement) against every known answer the reference's own
tests hold for this path (tests/golden/reference_fixtures.json)."""
import numpy as np
import pytest


def search_system(O, fx, criterion_nt=None):
    """Reference-layout system of test/Algorithm/NeighbourSearchTest.jl:12-65."""
    g = fx["neighbour_search"]
    L = O.make_layout(g["int_named_tuple"], g["float_named_tuple"], 9, basic_map=g["basic_index_map_dict"])
    IP = np.zeros((9, L.n_int), np.int32, order="F")
    FP = np.zeros((9, L.n_float), np.float32, order="F")
    for i, (x, y) in enumerate(g["positions"]):
        FP[i, L.PositionVec] = x
        FP[i, L.PositionVec + 1] = y
        IP[i, L.Tag] = g["tag"]
    d = O.domain2d(*g["domain"])
    adj = np.array([[1, 1], [1, 0]], np.int32)
    return O.System(IP, FP, np.ones(9, np.int32), np.zeros(9, np.int32), 9, L, d, adj, g["max_neighbour_number"])


def test_domain(oracle, golden):
    g = golden["domain"]
    d = oracle.domain2d(*g["args"])
    assert (d.n_x, d.n_y, d.n) == (g["n_x"], g["n_y"], g["n"])
    for k, v in g["approx"].items():
        assert getattr(d, k) == pytest.approx(v, rel=1e-6), k
    (x, y), (i, j) = g["cartesian_from_position"]
    import ctypes as C
    ci, cj = C.c_int32(), C.c_int32()
    oracle.lib().or_index_cartesian_from_position(C.byref(d), x, y, C.byref(ci), C.byref(cj))
    assert (ci.value, cj.value) == (i, j)
    assert oracle.lib().or_index_linear_from_position(C.byref(d), *g["linear_from_position"][0]) == \
        g["linear_from_position"][1]
    assert bool(oracle.lib().or_inside(C.byref(d), *g["inside"][0])) == g["inside"][1]


def test_stencil(oracle, golden):
    g = golden["stencil"]
    d = oracle.domain2d(*g["domain"])
    s = oracle.System(np.zeros((1, 1), np.int32), np.zeros((1, 2), np.float32), np.zeros(1, np.int32),
                      np.zeros(1, np.int32), 0, oracle.make_layout({"Tag": 1}, {"PositionVec": 2}, 1), d,
                      np.ones((1, 1), np.int32), 4)
    n = s.ns.n_cells
    assert (n, 9) == tuple(g["shape"])
    assert [s.ns.neighbour_cell_count[k] for k in range(n)] == g["count"]


def test_neighbour_search_fixture(oracle, golden):
    """nCount and sorted neighbour sets after search! then dynamic_search!."""
    g = golden["neighbour_search"]
    s = search_system(oracle, golden)
    assert s.search(oracle.CRIT_DEFAULT, 0) == 0
    assert s.search(oracle.CRIT_DEFAULT, 0) == 0  # dynamic_search! reuses the cached cells
    assert s.count().tolist() == g["n_count"]
    for i in range(9):
        assert sorted(s.neighbours(i).tolist()) == g["neighbours"][i]


def test_action_fixture(oracle, golden):
    """test/Algorithm/ActionTest.jl: self zeroes RVec, interaction adds 1 per neighbour."""
    s = search_system(oracle, golden)
    s.search(oracle.CRIT_DEFAULT, 0)
    col = s.L.PositionVec
    for _ in range(2):  # static then dynamic variants
        for c in (col, col + 1):
            s.selfaction(10, column=c, value=0.0)
        for c in (col, col + 1):
            s.interaction(10, column=c, value=1.0)
    exp = np.array(golden["action"]["expected_columns_1_and_2"], np.float32)
    assert np.array_equal(s.FP[:9, col], exp) and np.array_equal(s.FP[:9, col + 1], exp)


def test_layout_offsets(oracle, golden):
    g = golden["particle_system"]
    ii, n_int = oracle.named_index(g["int_named_tuple"])
    fi, n_float = oracle.named_index(g["float_named_tuple"])
    assert (n_int, n_float) == (g["n_int_capacity"], g["n_float_capacity"])
    got = {**ii, **fi}
    for k, v in g["basic_index"].items():
        assert got[k] == v, k
    gn = golden["named_index"]
    assert oracle.named_index(gn["int_named_tuple"])[1] == gn["n_int_capacity"]


def test_geometry_count(oracle, golden):
    g = golden["geometry"]
    pos, vol, gap = oracle.rectangle_create(g["gap"], *g["rectangle"])
    assert pos.shape[0] == g["count"]
    assert np.allclose(vol, g["gap"] ** 2, rtol=1e-6) and np.allclose(gap, g["gap"], rtol=1e-6)
