"""eppx_ps_download_noalloc (the download staged through the library's fixed
device scratch, for callers that cannot allocate) against eppx_ps_download
on the same particle systems: every table, is_alive and cell_index
identical, on two calls in a row (the second on a smaller system, so stale
scratch contents of the first would show), and oversized tables rejected
with EPPX_E_ARG before any device work."""
import ctypes as C

import numpy as np
import pytest

import parity_util as U

pytestmark = pytest.mark.gpu


def _download(ps, noalloc):
    from eppx import Class, _lib
    host = Class.HostParticleSystem(ps.n_capacity, ps.named_index_, ps.basic_parameters_, ps.basic_index_map_dict,
                                    ps.n_particles_)
    host.int_properties_[:] = -7  # every entry is written by the call
    host.float_properties_[:] = np.nan
    host.is_alive_[:] = -7
    host.cell_index_[:] = -7
    L = _lib.lib()
    fn = L.eppx_ps_download_noalloc if noalloc else L.eppx_ps_download
    n = C.c_int32()
    st = fn(ps.handle, ps.last_ns.handle if ps.last_ns is not None else None, C.byref(n),
            _lib.ptr(host.is_alive_, C.c_int32), _lib.ptr(host.cell_index_, C.c_int32),
            _lib.ptr(host.int_properties_, C.c_int32), _lib.ptr(host.float_properties_, C.c_float))
    return st, n.value, host


def _same(a, b):
    assert np.array_equal(a.is_alive_, b.is_alive_)
    assert np.array_equal(a.cell_index_, b.cell_index_)
    assert np.array_equal(a.int_properties_, b.int_properties_)
    # bitwise, NaN payloads included
    assert np.array_equal(a.float_properties_.view(np.uint32), b.float_properties_.view(np.uint32))


def _stepped(parallel, case, steps):
    from eppx import Algorithm, SPH
    ps, ns = U.device(parallel, case)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    if steps:
        Algorithm.step(ps, ns, case.physics, n_steps=steps)
    return ps, ns


def test_noalloc_download_matches_download_twice(parallel):
    from eppx import SPH, cases
    big = cases.small(M=40)            # 6,518 rows x 301 columns, neighbour columns exported
    small = cases.dam_break(0.04, 2.0, SPH.WendlandC2, M=64, fluid_width=0.5, fluid_height=0.6)
    assert small.n_particles < big.n_particles
    for case, steps in ((big, 7), (small, 3)):
        ps, ns = _stepped(parallel, case, steps)
        st0, n0, ref = _download(ps, noalloc=False)
        st1, n1, got = _download(ps, noalloc=True)
        assert st0 == 0 and st1 == 0
        assert n0 == n1 == case.n_particles
        _same(got, ref)
        # the neighbour columns came through too
        lay = ps.layout
        assert got.int_properties_[: case.n_particles, lay.nCount].sum() > 0
        ps.close()
        ns.close()


def test_noalloc_download_rejects_oversized_tables(parallel):
    from eppx import SPH, cases
    case = cases.dam_break(0.005, 2.0, SPH.WendlandC2, M=64)  # 86,018 rows x 469 columns > the scratch
    ps, ns = U.device(parallel, case)
    st, _, _ = _download(ps, noalloc=True)
    assert st == 1  # EPPX_E_ARG
    from eppx import _lib
    assert b"staging words" in _lib.lib().eppx_last_error(parallel.ctx)
    # the regular download still serves it
    st0, n0, _ = _download(ps, noalloc=False)
    assert st0 == 0 and n0 == case.n_particles
    ps.close()
    ns.close()
