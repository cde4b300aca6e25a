"""Helpers shared by the GPU parity tests: the same case on the device (through
the C ABI) and in the restatement (the checker)."""
import ctypes as C

import numpy as np

import oracle as O
from eppx import Algorithm, Class, SPH, _lib, cases

ADJ = np.array([[1, 1], [1, 0]], np.int32)
STATE_FLOAT = ("PositionVec", "VelocityVec", "dVelocityVec", "Mass", "Volume", "Density", "dDensity", "Pressure",
               "H")


def device(par, case, M=None):
    M = M or case.M
    ps = Class.ParticleSystem(par, case.domain, case.n_particles, case.int_nt, case.float_nt,
                              case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(par, case.domain, cases.ACTIVE_PAIR, max_neighbour_number=M)
    return ps, ns


def reference(case, M=None):
    M = M or case.M
    h = case.host
    L = O.make_layout(case.int_nt, case.float_nt, h.n_capacity)
    return O.System(h.int_properties_.copy(order="F"), h.float_properties_.copy(order="F"), h.is_alive_.copy(),
                    h.cell_index_.copy(), h.n_particles_, L, O.domain2d(*case.domain.args), ADJ, M)


def oracle_physics(physics):
    return O.physics(**physics.as_dict())


def search_both(par, case, M=None):
    ps, ns = device(par, case, M)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    ref = reference(case, M)
    assert ref.search(O.CRIT_SYMMETRY, 1, oracle_physics(case.physics)) == 0
    return ps, ns, ref


def download(ps):
    host = Class.mirror(ps)
    Class.syncto(host, ps)
    return host


def width(name):
    return 2 if name.endswith("Vec") else 1


def assert_state_equal(host, ref, names=STATE_FLOAT, alive_only=True):
    assert np.array_equal(host.is_alive_[: ref.n], ref.is_alive[: ref.n])
    rows = ref.is_alive[: ref.n] == 1 if alive_only else slice(0, ref.n)
    for name in names:
        c = getattr(ref.L, name)
        got = host.float_properties_[: ref.n][rows, c: c + width(name)]
        exp = ref.FP[: ref.n][rows, c: c + width(name)]
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            r, k = bad[0]
            raise AssertionError(f"{name}: {len(bad)} entries differ, first row {r}: {got[r, k]!r} vs {exp[r, k]!r}")


def assert_neighbour_columns_equal(host, ref):
    """nCount, nIndex in list order, nRVec, nR, nW, nDW, nHInv (alive rows)."""
    L = ref.L
    M = L.max_neighbour_number
    alive = np.nonzero(ref.is_alive[: ref.n] == 1)[0]
    cnt = ref.IP[alive, L.nCount]
    assert np.array_equal(host.int_properties_[alive, L.nCount], cnt)
    mask = np.arange(M)[None, :] < cnt[:, None]
    for col, w, table in (("nIndex", 1, "int"), ("nR", 1, "f"), ("nW", 1, "f"), ("nDW", 1, "f"),
                          ("nHInv", 1, "f"), ("nRVec", 2, "f")):
        c = getattr(L, col)
        got = (host.int_properties_ if table == "int" else host.float_properties_)[alive, c: c + w * M]
        exp = (ref.IP if table == "int" else ref.FP)[alive, c: c + w * M]
        m = np.repeat(mask, w, axis=1)
        if not np.array_equal(got[m], exp[m]):
            bad = np.argwhere((got != exp) & m)
            raise AssertionError(f"{col}: {len(bad)} entries differ (first {bad[0]})")


def export_pairs(ps, ns):
    L = _lib.lib()
    n = ps.n_particles_
    off = np.zeros(n + 1, np.int64)
    total = C.c_int64()
    _lib.check(L.eppx_pair_count(ps.handle, ns.handle, C.byref(total)), ps.parallel.ctx, "pair_count")
    js = np.zeros(max(1, total.value), np.int32)
    _lib.check(L.eppx_export_pairs(ps.handle, ns.handle, _lib.ptr(off, C.c_int64), _lib.ptr(js, C.c_int32),
                                   js.shape[0]), ps.parallel.ctx, "export_pairs")
    return off, js[: total.value]


def reference_pairs_in_order(ref):
    """CSR of the restatement's neighbour lists in list order (0-based)."""
    L = ref.L
    cnt = np.where(ref.is_alive[: ref.n] == 1, ref.IP[: ref.n, L.nCount], 0).astype(np.int64)
    off = np.zeros(ref.n + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    M = L.max_neighbour_number
    lists = ref.IP[: ref.n, L.nIndex: L.nIndex + M] - 1
    mask = np.arange(M)[None, :] < cnt[:, None]
    return off, lists[mask].astype(np.int32)
