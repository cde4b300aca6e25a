"""Host logic of the serving loop (eppx.Serve) on the CPU, with a stand-in
batch and a virtual clock: first-come-first-served admission by the HBM
budget and case slots, chunked steps, every request served once, time to
completion."""
import numpy as np
import pytest

from eppx import Serve


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class Host:
    def __init__(self, n):
        self.n_particles_ = n


class FakeBatch:
    """Stand-in for eppx.Batch: each batch step advances every case with steps
    left by one, costs `cost` seconds of the clock and adds n pairs per
    particle of each advancing case."""

    def __init__(self, clock, case_capacity, cost=1e-3):
        self.clock, self.cap, self.cost = clock, case_capacity, cost
        self.cases = {}
        self.pairs = 0
        self.log = []

    def add(self, domain, physics, host, steps):
        cid = min(set(range(self.cap)) - set(self.cases))
        self.cases[cid] = [steps, host]
        self.log.append(("add", cid, host.n_particles_))
        return cid

    def steps_left(self):
        out = np.full(self.cap, -1, np.int32)
        for c, (left, _) in self.cases.items():
            out[c] = left
        return out

    def step(self, n, sync=True):
        self.log.append(("step", n))
        for _ in range(n):
            for v in self.cases.values():
                if v[0] > 0:
                    v[0] -= 1
                    self.pairs += 10 * v[1].n_particles_
            self.clock.t += self.cost

    def pair_total(self, reset=False):
        v = self.pairs
        if reset:
            self.pairs = 0
        return v

    def download(self, cid):
        return self.cases[cid][1]

    def remove(self, cid):
        del self.cases[cid]

    def close(self):
        pass


def serve(reqs, slots, case_slots, chunk=10, cost=1e-3):
    clock = Clock()
    fb = FakeBatch(clock, case_slots, cost)
    s = Serve.Server(None, None, slots * 100, case_slots, None, chunk=chunk, batch=fb, bytes_per_slot=100)
    res, st = s.run(reqs, clock=clock, sleep=clock.sleep)
    return res, st, fb


def R(t, n, steps):
    return Serve.Request(t, None, None, Host(n), steps)


def test_every_request_is_served_once_in_stream_order():
    reqs = [R(0.0, 100, 30), R(0.005, 50, 5), R(0.006, 70, 12), R(0.2, 10, 3)]
    res, st, fb = serve(reqs, slots=1000, case_slots=8)
    assert [r.index for r in res] == [0, 1, 2, 3]
    assert all(r.host is q.host for r, q in zip(res, reqs))
    for r, q in zip(res, reqs):
        assert r.admitted >= q.arrival and r.finished >= r.admitted + q.steps * 1e-3 * 0.999
    assert st.pairs == sum(10 * q.n_particles * q.steps for q in reqs)
    assert st.admitted_order == [0, 1, 2, 3]
    s = Serve.summary(res, st)
    assert s["p50"] <= s["p99"] <= s["max"] and s["pair_interactions_per_s"] > 0


def test_admission_waits_for_budget_and_case_slots():
    # the budget holds 150 particles: request 1 (100) must wait until request 0 (100) is done
    reqs = [R(0.0, 100, 20), R(0.001, 100, 5), R(0.002, 40, 5)]
    res, st, fb = serve(reqs, slots=150, case_slots=8)
    assert st.peak_slots <= 150
    assert res[1].admitted >= res[0].finished - 1e-12
    # first come, first served: request 2 fits beside request 0 but does not overtake request 1
    assert st.admitted_order == [0, 1, 2]
    # case slots: with one slot the cases run one after another
    res, st, fb = serve([R(0.0, 10, 4), R(0.0, 10, 4)], slots=1000, case_slots=1)
    assert st.peak_cases == 1 and res[1].admitted >= res[0].finished - 1e-12


def test_chunks_let_arrivals_join_running_cases():
    # a long case is running when a short one arrives: the short one joins at the next chunk boundary
    reqs = [R(0.0, 100, 200), R(0.015, 10, 5)]
    res, st, fb = serve(reqs, slots=1000, case_slots=4, chunk=10, cost=1e-3)
    assert res[1].admitted < res[0].finished
    assert res[1].finished < res[0].finished
    assert max(n for kind, *rest in fb.log if kind == "step" for n in rest) <= 10


def test_bad_streams_are_rejected():
    with pytest.raises(ValueError):
        serve([R(0.0, 2000, 3)], slots=1000, case_slots=2)
    with pytest.raises(ValueError):
        serve([R(0.5, 10, 3), R(0.1, 10, 3)], slots=1000, case_slots=2)


def test_percentile():
    assert Serve.percentile([], 0.5) is None
    assert Serve.percentile([3.0, 1.0, 2.0], 0.5) == 2.0
    assert Serve.percentile([0.0, 10.0], 0.99) == pytest.approx(9.9)
