"""A case driven by a configuration and a raw file (main.jl:19-48 with
DataIO's Input.jl and Writer.jl) runs bit-identically to the same case built
directly, and to the restatement."""
import numpy as np
import pytest

import oracle as O
import parity_util as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from eppx import cases
    return cases.small()


def test_config_driven_case_matches_direct(parallel, small, tmp_path):
    from eppx import Algorithm, DataIO, SPH
    ps_d, ns_d = U.device(parallel, small)
    w = DataIO.Writer(str(tmp_path))
    c = DataIO.config(parallel, small.domain, ps_d, ns_d, w)
    app = DataIO.appendix()
    raw = DataIO.save(w, small.host, app)
    c["particle_system"]["load"] = {"mode": "continue", "file": raw}
    c2 = DataIO.parse(DataIO.writeConfig(w, c))
    assert DataIO.Parallel(c2).device == parallel.device
    domain = DataIO.Domain(c2)
    ps_c = DataIO.ParticleSystem(c2, parallel, domain)
    ns_c = DataIO.NeighbourSystem(c2, parallel, domain)
    assert ps_c.n_particles_ == small.n_particles and ps_c.n_capacity == ps_d.n_capacity
    n = 25
    for ps, ns, d in ((ps_d, ns_d, small.domain), (ps_c, ns_c, domain)):
        Algorithm.search(ps, d, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(small.physics))
        Algorithm.step(ps, ns, small.physics, n_steps=n)
    a, b = U.download(ps_d), U.download(ps_c)
    assert np.array_equal(a.is_alive_, b.is_alive_)
    assert np.array_equal(a.cell_index_, b.cell_index_)
    assert np.array_equal(a.int_properties_, b.int_properties_)
    assert np.array_equal(a.float_properties_.view(np.uint32), b.float_properties_.view(np.uint32))
    ref = U.reference(small)
    assert ref.search(O.CRIT_SYMMETRY, 1, U.oracle_physics(small.physics)) == 0
    err, _ = ref.run_steps(n, U.oracle_physics(small.physics))
    assert err == 0
    U.assert_state_equal(b, ref)


@pytest.mark.parametrize("physics", ["benchmark", "example"])
def test_collapse_dry_driver_matches_direct_run(parallel, tmp_path, physics):
    """examples/collapse_dry.py (casegen, then main with the filter every 20
    steps and an output every 40), with the benchmark's constants or
    main.jl's own (linear EOS, classic viscosity): the written state equals a
    direct run."""
    import importlib.util
    import os
    from eppx import Algorithm, DataIO, SPH, cases
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "etherparallelparticles.jl_amd",
                        "examples", "collapse_dry.py")
    spec = importlib.util.spec_from_file_location("collapse_dry", path)
    ex = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ex)
    out = str(tmp_path / "run")
    ex.casegen(out, dx=0.02, physics=physics)
    app = ex.main(out, 60, filter_interval=20, output_interval=40, log=lambda *_: None)
    assert app["TMSTEP"] == 40 and app["WriteStep"] == 1  # the output of step 40
    case = cases.dam_break(0.02, 1.5, SPH.CubicSpline, M=40)
    ph = SPH.example_physics(0.02) if physics == "example" else case.physics
    ps, ns = U.device(parallel, case)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(ph))
    Algorithm.step(ps, ns, ph, n_steps=40, filter_interval=20)
    direct = U.download(ps)
    w = DataIO.WriterFrom(DataIO.parse(os.path.join(out, "config", "config.json")))
    saved = Class_host_like(direct)
    DataIO.load_into(w, saved, 1)
    keep = direct.is_alive_[: direct.n_particles_] == 1
    assert np.array_equal(saved.float_properties_[: saved.n_particles_].view(np.uint32),
                          direct.float_properties_[: direct.n_particles_][keep].view(np.uint32))
    assert np.array_equal(saved.int_properties_[: saved.n_particles_],
                          direct.int_properties_[: direct.n_particles_][keep])


def Class_host_like(h):
    from eppx import Class
    return Class.HostParticleSystem(h.n_capacity, h.named_index_, h.basic_parameters_)
