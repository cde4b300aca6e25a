"""DataIO (src/DataIO/Config/Input.jl, Output.jl, Writer/RawWriter.jl): the
configuration dict to handles and back, the capacity_expand lambda without
eval, and raw output round trips.  Host only."""
import json
import os
import types
import warnings

import numpy as np
import pytest

from eppx import Class, DataIO, cases


def _reference_style_config():
    """A configuration as the reference's template lays it out (section and
    key names of template.json), with its values."""
    return {
        "parallel": {"int": "Int32", "float": "Float32", "backend": "cpu", "device": 1},
        "domain": {"dimension": 2, "gap": 0.15, "first_x": 0.1, "first_y": 0.2, "first_z": 0.0, "last_x": 0.9,
                   "last_y": 0.8, "last_z": 0.0},
        "particle_system": {
            "n_particles": 100, "capacity_expand": "n -> n + 100", "load": {"mode": "initial", "file": "demo.jld2"},
            "int_named_tuple": {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": 40},
            "float_named_tuple": {"PositionVec": 2, "VelocityVec": 2, "dVelocityVec": 2, "AccelerationVec": 2,
                                  "Mass": 1, "Density": 1, "dDensity": 1, "Volume": 1, "Pressure": 1, "Gap": 1,
                                  "H": 1, "SumWeight": 1, "SumWeightedDensity": 1, "SumWeightedPressure": 1,
                                  "nW": 40, "nDW": 40, "nHInv": 40, "nRVec": 80, "nR": 40},
            "basic_parameters": {"c_0": 340.0, "gamma": 1.4, "mu": 1e-3},
            "basic_index_map_dict": {"PositionVec": "PositionVec", "Tag": "Tag", "nCount": "nCount",
                                     "nIndex": "nIndex", "nRVec": "nRVec", "nR": "nR"}},
        "neighbour_system": {"active_pair": [[1, 1], [1, 2], [2, 1]],
                             "periodic_boundary_policy": {"type": "none", "axis": [False, False, False]},
                             "max_neighbour_number": 50, "n_threads": 256},
        "writer": {"path": "demo", "file_name": "result", "connect": "_", "digits": 4, "suffix": ".jld2"},
    }


def test_template_has_the_reference_sections():
    t, r = DataIO.template(), _reference_style_config()
    assert list(t) == list(r)
    for sec in r:
        assert set(t[sec]) == set(r[sec]), sec
    assert set(t["particle_system"]["load"]) == {"mode", "file"}


@pytest.mark.parametrize("src,n,cap", [
    ("n -> n + 100", 100, 200),       # what Output.jl:57 writes
    ("n -> n", 7, 7),
    ("n->2n + 1", 10, 21),            # Julia's juxtaposed coefficient
    ("n -> n + div(n, 10)", 95, 104),
    ("n -> n ÷ 2 + n", 9, 13),
    ("n -> (n + 3) * 2 - n", 5, 11),
    ("n -> 2n + (1 - 2n) ÷ 4", 4, 7),  # -7 ÷ 4 truncates toward zero: -1 (floor division gives -2)
    ("n -> n + (n - 27021597764222980) ÷ 3 + 9007199254740993", 1, 1),  # -(2^53 + 1) exactly, past float64
])
def test_capacity_expand(src, n, cap):
    assert DataIO.capacity_expand(src, n) == cap


@pytest.mark.parametrize("src", ["n -> run(`ls`)", "n -> n + 1.5", "x -> x", "n -> n - 1", "n -> __import__('os')",
                                 "n -> n ÷ 0", "n -> n ** 2", "n -> n // 2 + n"])
def test_capacity_expand_refuses(src):
    with pytest.raises(ValueError):
        DataIO.capacity_expand(src, 4)


def test_parallel_from_config():
    c = _reference_style_config()
    p = DataIO.Parallel(c)
    assert p.device == 0  # the reference numbers devices from 1
    c["parallel"]["int"] = "Int64"
    with pytest.raises(NotImplementedError):
        DataIO.Parallel(c)
    c = _reference_style_config()
    c["parallel"]["backend"] = "tpu"
    with pytest.raises(KeyError):
        DataIO.Parallel(c)


def test_domain_from_config_is_domain2d():
    c = _reference_style_config()
    d = DataIO.Domain(c)
    e = Class.Domain2D(0.15, 0.1, 0.2, 0.9, 0.8)
    for f, _ in type(d.c)._fields_:
        assert getattr(d.c, f) == getattr(e.c, f), f
    c["domain"]["dimension"] = 3
    with pytest.raises(NotImplementedError):
        DataIO.Domain(c)


def test_continue_from_jld2_is_refused():
    c = _reference_style_config()
    c["particle_system"]["load"]["mode"] = "continue"
    with pytest.raises(NotImplementedError):
        DataIO.ParticleSystem(c, DataIO.Parallel(c), DataIO.Domain(c))


def test_raw_save_load_round_trip(tmp_path):
    case = cases.small()
    host = case.host
    host.is_alive_[[3, 17, 400]] = 0  # dead rows are not written (RawWriter.jl:40-41)
    w = DataIO.Writer(str(tmp_path / "out"), digits=5)
    app = DataIO.appendix()
    app.update(TMSTEP=400, TimeValue=0.25, WriteStep=3)
    name = DataIO.save(w, host, app)
    assert name == os.path.join(str(tmp_path / "out"), "raw", "result_00003.npz")
    back = Class.HostParticleSystem(host.n_capacity, host.named_index_, host.basic_parameters_)
    got = DataIO.load_into(w, back, 3)
    keep = host.is_alive_[: host.n_particles_] == 1
    n = int(keep.sum())
    assert back.n_particles_ == n and (back.is_alive_[:n] == 1).all() and (back.is_alive_[n:] == 0).all()
    assert np.array_equal(back.int_properties_[:n], host.int_properties_[: host.n_particles_][keep])
    assert np.array_equal(back.float_properties_[:n], host.float_properties_[: host.n_particles_][keep])
    assert got["TMSTEP"] == 400 and got["TimeValue"] == 0.25 and got["WriteStep"] == 3
    assert len(got["TimeStamp"]) == len("yyyy_mm_dd_HH_MM_SS")


def test_jld2_writer_writes_npz(tmp_path):
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        w = DataIO.WriterFrom(dict(_reference_style_config(), writer={
            "path": str(tmp_path), "file_name": "r", "connect": "-", "digits": 2, "suffix": ".jld2"}))
    assert any("HDF5" in str(x.message) for x in caught)
    assert w.get_file_name(7).endswith(os.path.join("raw", "r-07.npz"))


def test_config_round_trip(tmp_path):
    """Output then Input gives back the same domain, capacity, layout and pairs."""
    case = cases.small()
    host = case.host
    par = DataIO.Parallel(DataIO.template())
    ns = types.SimpleNamespace(active_pair_=Class.ActivePair(cases.ACTIVE_PAIR), max_neighbour_number=40,
                               periodic_boundary_=Class.PeriodicBoundary(par, case.domain))
    w = DataIO.Writer(str(tmp_path))
    c = DataIO.config(par, case.domain, host, ns, w)
    name = DataIO.writeConfig(w, c)
    c2 = DataIO.parse(name)
    assert c2 == json.loads(json.dumps(c))
    d = DataIO.Domain(c2)
    for f, _ in type(d.c)._fields_:
        assert getattr(d.c, f) == getattr(case.domain.c, f), f
    ps = c2["particle_system"]
    assert DataIO.capacity_expand(ps["capacity_expand"], ps["n_particles"]) == host.n_capacity
    ni = Class.NamedIndex(ps["int_named_tuple"], ps["float_named_tuple"])
    assert ni.get_index_named_tuple() == host.named_index_.get_index_named_tuple()
    assert [tuple(p) for p in c2["neighbour_system"]["active_pair"]] == list(cases.ACTIVE_PAIR)
    assert ps["basic_parameters"]["rho0"] == case.basic_parameters["rho0"]
    import yaml
    y = yaml.safe_load(open(DataIO.writeConfig(w, c, "yaml")))
    assert y == json.loads(json.dumps(c))


def _example():
    import importlib.util
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "etherparallelparticles.jl_amd",
                        "examples", "collapse_dry.py")
    spec = importlib.util.spec_from_file_location("collapse_dry", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_casegen_writes_a_loadable_case(tmp_path):
    """casegen.jl's output: a configuration and write step 0 that load back
    into the same tables."""
    ex = _example()
    name = ex.casegen(str(tmp_path / "case"), dx=0.02)
    c = DataIO.parse(name)
    case = cases.dam_break(0.02, 1.5, ex.KERNELS["cubic"], M=40)
    assert c["particle_system"]["n_particles"] == case.n_particles
    w = DataIO.WriterFrom(c)
    back = Class.HostParticleSystem(case.host.n_capacity, case.host.named_index_, case.host.basic_parameters_)
    app = DataIO.load_into(w, back, 0)
    assert app["WriteStep"] == 0 and back.n_particles_ == case.n_particles
    assert np.array_equal(back.float_properties_, case.host.float_properties_)
    assert np.array_equal(back.int_properties_, case.host.int_properties_)


def test_casegen_records_the_driver_physics(tmp_path):
    """--physics example: main.jl's constants (linear EOS, c0 = 120, classic
    viscosity mu0 = 1e-3, g = 9.8, dt = 0.2 h0 / c0) reach the driver."""
    from eppx import SPH, _lib
    ex = _example()
    c = DataIO.parse(ex.casegen(str(tmp_path / "case"), dx=0.02, physics="example"))
    ph = ex.driver_physics(c["driver"])
    assert ph.eos == _lib.EOS_LINEAR and ph.visc == _lib.VISC_CLASSIC
    assert np.float32(ph.c0) == np.float32(120.0) and np.float32(ph.gy) == np.float32(-9.8)
    assert np.float32(ph.mu) == np.float32(1e-3)
    assert np.float32(ph.dt) == np.float32(np.float32(0.1 * np.float64(np.float32(0.03)) / 120.0) * 2)
    base = ex.driver_physics(DataIO.parse(ex.casegen(str(tmp_path / "b"), dx=0.02))["driver"])
    assert base.eos == _lib.EOS_TAIT and base.visc == _lib.VISC_ARTIFICIAL
