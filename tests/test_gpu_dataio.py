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
