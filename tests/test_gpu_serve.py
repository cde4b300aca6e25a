"""The serving loop (eppx.Serve) on the device: a short Poisson stream of
dam-break requests (seed 1) served by continuous batching -- requests join
the running batch between chunks and leave when their budgets are done --
and each result equals the same case run alone on the particle-system path
(FAST math; identical bits: a case's summation order is its own, whatever
shares its steps)."""
import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(parallel):
    from eppx import Algorithm, Batch, Class, SPH, Serve, cases
    reqs = cases.serving_requests(1, n_requests=6, rate_hz=400.0, n_min=1.5e3, n_max=1.2e4, steps_min=20,
                                  steps_max=70, M=64)
    biggest = max(r.n_particles for r in reqs)
    # room for about two of the larger cases at once: later arrivals wait for slots
    budget = 2 * biggest * Batch.bytes_per_particle()
    parallel.set_math("fast")
    try:
        srv = Serve.Server(parallel, reqs[0].host, budget, case_slots=4, active_pair=cases.ACTIVE_PAIR,
                           max_neighbour_number=64, chunk=16)
        try:
            results, stats = srv.run(reqs)
        finally:
            srv.close()
        s = Serve.summary(results, stats)
        print(f"served {len(reqs)} requests: p50 {s['p50']:.4f} s, p99 {s['p99']:.4f} s, {stats.chunks} chunks, "
              f"peak {stats.peak_cases} cases / {stats.peak_slots} slots, {stats.pairs} pairs")
        assert all(r is not None for r in results)
        assert stats.peak_slots <= srv.slots and stats.peak_cases >= 2  # cases overlapped
        for req, res in zip(reqs, results):
            ni = req.host.named_index_
            c_ps = Class.ParticleSystem(parallel, req.domain, req.n_particles, dict(ni.int_table.capacity),
                                        dict(ni.float_table.capacity), req.host.basic_parameters_)
            Class.syncto(c_ps, req.host)
            ns = Class.NeighbourSystem(parallel, req.domain, cases.ACTIVE_PAIR, max_neighbour_number=64)
            Algorithm.search(c_ps, req.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                             action=SPH.sComputeKernelAndiContinuity(req.physics))
            Algorithm.step(c_ps, ns, req.physics, n_steps=req.steps)
            alone = U.download(c_ps)
            n = req.n_particles
            assert np.array_equal(res.host.is_alive_[:n], alone.is_alive_[:n])
            L = c_ps.layout
            for name in U.STATE_FLOAT:
                c = getattr(L, name)
                w = U.width(name)
                assert np.array_equal(res.host.float_properties_[:n, c: c + w].view(np.uint32),
                                      alone.float_properties_[:n, c: c + w].view(np.uint32)), (req.name, name)
            c_ps.close()
            ns.close()
    finally:
        parallel.set_math("exact")
