"""The reference's own test fixtures, run through libeppx (C ABI) on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def search_fixture(par, golden):
    from eppx import Class
    g = golden["neighbour_search"]
    domain = Class.Domain2D(*g["domain"])
    ps = Class.ParticleSystem(par, domain, 9, g["int_named_tuple"], g["float_named_tuple"],
                              {"c_0": 340.0, "gamma": 7, "mu": 1e-3},
                              basic_index_map_dict=g["basic_index_map_dict"])
    host = Class.mirror(ps)
    bi = host.basic_index_
    for i, (x, y) in enumerate(g["positions"]):
        host.float_properties_[i, bi["PositionVec"] - 1] = x
        host.float_properties_[i, bi["PositionVec"]] = y
        host.is_alive_[i] = 1
        host.int_properties_[i, bi["Tag"] - 1] = 1
    Class.asyncto(ps, host)
    ns = Class.NeighbourSystem(par, domain, [tuple(p) for p in g["active_pair"]])
    return ps, ns, domain, host


def test_neighbour_search_fixture(parallel, golden):
    """test/Algorithm/NeighbourSearchTest.jl:66-152 through the C ABI."""
    from eppx import Algorithm, Class
    g = golden["neighbour_search"]
    ps, ns, domain, host = search_fixture(parallel, golden)
    Algorithm.search(ps, domain, ns)
    Algorithm.dynamic_search(ps, domain, ns)
    Class.asyncto(host, ps)
    bi = host.basic_index_
    counts = host.int_properties_[:, bi["nCount"] - 1]
    assert counts.tolist() == g["n_count"]
    for i in range(9):
        lst = host.int_properties_[i, bi["nIndex"] - 1: bi["nIndex"] - 1 + counts[i]]
        assert sorted(lst.tolist()) == g["neighbours"][i]


def test_search_fixture_columns_match_restatement(parallel, golden, oracle):
    """nIndex list order, nRVec and nR bit-identical to the restatement."""
    from eppx import Algorithm, Class
    from test_oracle_golden import search_system
    ps, ns, domain, host = search_fixture(parallel, golden)
    Algorithm.search(ps, domain, ns)
    Class.asyncto(host, ps)
    ref = search_system(oracle, golden)
    ref.search(oracle.CRIT_DEFAULT, 0)
    L = ref.L
    for col, w in (("nIndex", 1), ("nRVec", 2), ("nR", 1)):
        c = getattr(L, col)
        for i in range(9):
            k = int(ref.IP[i, L.nCount]) * w
            table_h = host.int_properties_ if col == "nIndex" else host.float_properties_
            table_r = ref.IP if col == "nIndex" else ref.FP
            assert np.array_equal(table_h[i, c: c + k], table_r[i, c: c + k]), (col, i)


def test_action_fixture(parallel, golden):
    """test/Algorithm/ActionTest.jl:219-239: self zeroes RVec, interaction adds 1 per neighbour."""
    from eppx import Algorithm, Class, SPH
    ps, ns, domain, host = search_fixture(parallel, golden)
    Algorithm.search(ps, domain, ns)
    for sa, ia in ((Algorithm.selfaction, Algorithm.interaction),
                   (Algorithm.dynamic_selfaction, Algorithm.dynamic_interaction)):
        sa(ps, SPH.ScaleField("RVec", 0, 0.0))
        sa(ps, SPH.ScaleField("RVec", 1, 0.0))
        ia(ps, SPH.AddPerNeighbour("RVec", 0, 1.0))
        ia(ps, SPH.AddPerNeighbour("RVec", 1, 1.0))
    Class.asyncto(host, ps)
    exp = np.array(golden["action"]["expected_columns_1_and_2"], np.float32)
    assert np.array_equal(host.float_properties_[:9, 0], exp)
    assert np.array_equal(host.float_properties_[:9, 1], exp)


def test_cell_stencil(parallel, golden, oracle):
    """device_initializeNeighbourSystem! table (NeighbourSystemBaseTest.jl:20-24)."""
    from eppx import Class
    g = golden["stencil"]
    domain = Class.Domain2D(*g["domain"])
    ns = Class.NeighbourSystem(parallel, domain, [(1, 1), (1, 2), (2, 1)])
    count, lst = ns.neighbour_cell_index()
    assert count.tolist() == g["count"] and lst.shape == tuple(g["shape"])
    d = oracle.domain2d(*g["domain"])
    s = oracle.System(np.zeros((1, 1), np.int32), np.zeros((1, 2), np.float32), np.zeros(1, np.int32),
                      np.zeros(1, np.int32), 0, oracle.make_layout({"Tag": 1}, {"PositionVec": 2}, 1), d,
                      np.ones((1, 1), np.int32), 4)
    n = s.ns.n_cells
    ref = np.array([[s.ns.neighbour_cell_list[i + n * k] for k in range(9)] for i in range(n)])
    assert np.array_equal(lst, ref)
