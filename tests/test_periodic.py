"""Periodic boundary policies (src/Class/NeighbourSystem/PeriodicBoundary.jl,
DataIO Input.jl:84-92 / Output.jl:86-97): the policy types, the shift table's
shapes from the reference's own PeriodicBoundaryTest.jl, and the policy through
a configuration and back.  Host only."""
import json
import types

import numpy as np
import pytest

from eppx import Class, DataIO, Environment, cases


def test_shift_table_shapes_golden():
    """PeriodicBoundaryTest.jl:11-31: Domain2D(0.15, 0.1, 0.2, 0.9, 0.9) has 5 x 4
    cells; NonePeriodicBoundaryPolicy gives a (1, 1, 1) table, AlongX (20, 9, 2)."""
    par = Environment.Parallel()  # the table is host data: no device context is made
    d = Class.Domain2D(0.15, 0.1, 0.2, 0.9, 0.9)
    assert d.get_n() == 20
    pb = Class.PeriodicBoundary(par, d, Class.NonePeriodicBoundaryPolicy)
    assert pb.neighbour_cell_relative_position_list_.shape == (1, 1, 1)
    pb = Class.PeriodicBoundary(par, d, Class.PeriodicBoundaryPolicy2DAlongX)
    assert pb.neighbour_cell_relative_position_list_.shape == (20, 9, 2)


def test_shift_table_3d_shape_and_zero():
    """PeriodicBoundary.jl:48-60 in 3D: (n_cells, 27, 3), zero-filled (the TODO at :57)
    and kept read-only, so the shift added at NeighbourSearch.jl:129-139 is zero."""
    d = Class.Domain3D(0.1, 0.0, 0.0, 0.0, 0.5, 0.3, 0.2)
    pb = Class.PeriodicBoundary(None, d, Class.PeriodicBoundaryPolicy3DAlongXYZ)
    t = pb.neighbour_cell_relative_position_list_
    assert t.shape == (d.get_n(), 27, 3) and t.dtype == np.float32
    assert not t.any()
    with pytest.raises(ValueError):
        t[0, 0, 0] = 1.0
    assert "PeriodicBoundaryPolicy3D{true, true, true}" in repr(pb)


def test_named_policies():
    """The aliases of PeriodicBoundary.jl:17-26 carry their axes; a policy type
    is built once per axis set, like the reference's parametric types."""
    P2, P3 = Class.PeriodicBoundaryPolicy2D, Class.PeriodicBoundaryPolicy3D
    assert Class.PeriodicBoundaryPolicy2DAlongX.axis == (True, False)
    assert Class.PeriodicBoundaryPolicy2DAlongY.axis == (False, True)
    assert Class.PeriodicBoundaryPolicy2DAlongXY.axis == (True, True)
    assert Class.PeriodicBoundaryPolicy3DAlongZX.axis == (True, False, True)
    assert Class.PeriodicBoundaryPolicy3DAlongYZ.axis == (False, True, True)
    assert P2.of(True, False) is Class.PeriodicBoundaryPolicy2DAlongX
    assert P3.of(False, False, True) is Class.PeriodicBoundaryPolicy3DAlongZ
    for p in (Class.PeriodicBoundaryPolicy2DAlongXY, Class.PeriodicBoundaryPolicy3DAlongX):
        assert issubclass(p, Class.HavePeriodicBoundaryPolicy)
        assert issubclass(p, Class.AbstractPeriodicBoundaryPolicy)
    assert not issubclass(Class.NonePeriodicBoundaryPolicy, Class.HavePeriodicBoundaryPolicy)
    d = Class.Domain2D(0.15, 0.1, 0.2, 0.9, 0.9)
    for bad in (int, Class.HavePeriodicBoundaryPolicy, Class.PeriodicBoundaryPolicy2D, "none"):
        with pytest.raises(TypeError):
            Class.PeriodicBoundary(None, d, bad)
    with pytest.raises(TypeError):
        P2.of(True, False, True)


@pytest.mark.parametrize("section, dim, expected", [
    ({"type": "none", "axis": [True, True, True]}, 2, Class.NonePeriodicBoundaryPolicy),
    ({"type": "2D", "axis": [True, False, False]}, 2, Class.PeriodicBoundaryPolicy2DAlongX),
    ({"type": "2D", "axis": [False, True, True]}, 2, Class.PeriodicBoundaryPolicy2DAlongY),
    ({"type": "3D", "axis": [True, False, True]}, 3, Class.PeriodicBoundaryPolicy3DAlongZX),
])
def test_policy_from_config(section, dim, expected):
    """Input.jl:84-92: "none", else the policy of the domain's dimension from the
    first N axis flags (the section's type string is not consulted)."""
    d = Class.Domain2D(0.15, 0.1, 0.2, 0.9, 0.9) if dim == 2 else Class.Domain3D(0.1, 0, 0, 0, 0.5, 0.3, 0.2)
    assert DataIO.periodic_policy(section, d) is expected


def test_policy_from_config_refuses_non_boolean_axis():
    d = Class.Domain2D(0.15, 0.1, 0.2, 0.9, 0.9)
    with pytest.raises(ValueError):
        DataIO.periodic_policy({"type": "2D", "axis": [1, 0, 0]}, d)


@pytest.mark.parametrize("policy, section", [
    (Class.NonePeriodicBoundaryPolicy, {"type": "none", "axis": [False, False, False]}),
    (Class.PeriodicBoundaryPolicy2DAlongY, {"type": "2D", "axis": [False, True, False]}),
    (Class.PeriodicBoundaryPolicy2DAlongXY, {"type": "2D", "axis": [True, True, False]}),
])
def test_policy_config_round_trip(tmp_path, policy, section):
    """Output.jl:86-97 writes the policy's type and axis flags; Input reads the
    same policy back."""
    case = cases.small()
    par = DataIO.Parallel(DataIO.template())
    ns = types.SimpleNamespace(active_pair_=Class.ActivePair(cases.ACTIVE_PAIR), max_neighbour_number=40,
                               periodic_boundary_=Class.PeriodicBoundary(par, case.domain, policy))
    w = DataIO.Writer(str(tmp_path))
    c = DataIO.config(par, case.domain, case.host, ns, w)
    assert c["neighbour_system"]["periodic_boundary_policy"] == section
    c2 = DataIO.parse(DataIO.writeConfig(w, c))
    assert c2 == json.loads(json.dumps(c))
    assert DataIO.periodic_policy(c2["neighbour_system"]["periodic_boundary_policy"], DataIO.Domain(c2)) is policy
