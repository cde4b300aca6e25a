"""Serve: a request loop over one Batch (continuous batching of dam-break cases).

The reference has no serving path: its driver runs one case per process
(example/collapse_dry/same/main.jl:148-207).  Here requests arrive over
time, each a dam-break case (its host tables, domain, physics and step
budget).  The loop admits a request as soon as its particles fit the HBM
budget (admission by device bytes per particle slot, Batch.bytes_per_particle)
and a case slot is free, keeps every admitted case resident in the batch,
advances all of them together in chunks of at most `chunk` batch steps (so
new arrivals join between chunks while the others keep running), and
returns a case -- downloaded, its slots released -- as soon as its budget is
done.  A case's result does not depend on which other cases shared its
steps (the per-case summation order is the case's own).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from .Batch import Batch, bytes_per_particle


@dataclass
class Request:
    arrival: float                # seconds after the start of the stream
    domain: object                # Class.Domain2D
    physics: object               # SPH.Physics
    host: object                  # Class.HostParticleSystem (its n_particles_ rows are the case)
    steps: int                    # step budget
    name: str = ""

    @property
    def n_particles(self) -> int:
        return self.host.n_particles_


@dataclass
class Result:
    index: int                    # position of the request in the stream
    host: object                  # the case's state after its budget
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        """Time to completion: arrival to result on the host."""
        return self.finished - self.arrival


@dataclass
class Stats:
    makespan: float = 0.0
    pairs: int = 0                # sum of nCount over every search of the served steps
    chunks: int = 0
    admitted_order: list = field(default_factory=list)
    peak_slots: int = 0
    peak_cases: int = 0

    def rate(self) -> float:
        return self.pairs / self.makespan if self.makespan > 0 else 0.0


def percentile(xs, q: float):
    """Linear-interpolated percentile (q in [0, 1]) of a list, None when empty."""
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


class Server:
    """Server(parallel, template, budget_bytes, case_slots, active_pair, max_neighbour_number, chunk).

    budget_bytes: HBM admitted to resident cases; the slot capacity is
    budget_bytes // Batch.bytes_per_particle().  `batch` replaces the device
    batch (an object with add / step / steps_left / download / remove /
    pair_total / close), which the host-logic tests use.
    """

    def __init__(self, parallel, template, budget_bytes: int, case_slots: int, active_pair,
                 max_neighbour_number: int = 64, chunk: int = 100, batch=None, bytes_per_slot: int | None = None):
        self.bytes_per_slot = bytes_per_slot if bytes_per_slot is not None else bytes_per_particle()
        self.slots = int(budget_bytes) // self.bytes_per_slot
        self.case_slots = int(case_slots)
        self.chunk = int(chunk)
        self.parallel = parallel
        self.batch = batch if batch is not None else Batch(parallel, template, self.slots, self.case_slots,
                                                           active_pair, max_neighbour_number)

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
        self.batch = None

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        """Serves the requests (sorted by arrival).  Returns (results in stream
        order, Stats).  A request larger than the whole budget raises
        ValueError before anything runs."""
        reqs = list(requests)
        for k, r in enumerate(reqs):
            if r.n_particles > self.slots:
                raise ValueError(f"request {k} ({r.n_particles} particles) exceeds the HBM budget "
                                 f"({self.slots} slots of {self.bytes_per_slot} B)")
            if k and r.arrival < reqs[k - 1].arrival:
                raise ValueError("requests must be sorted by arrival time")
        b = self.batch
        pending = list(range(len(reqs)))
        active = {}            # case id -> request index
        admitted_at = {}
        used = 0
        results = [None] * len(reqs)
        st = Stats()
        t0 = clock()
        while pending or active:
            now = clock() - t0
            # admit in arrival order while the next one fits (no overtaking: first come, first served)
            while pending and reqs[pending[0]].arrival <= now:
                k = pending[0]
                r = reqs[k]
                if used + r.n_particles > self.slots or len(active) >= self.case_slots:
                    break
                cid = b.add(r.domain, r.physics, r.host, r.steps)
                active[cid] = k
                admitted_at[k] = clock() - t0
                used += r.n_particles
                pending.pop(0)
                st.admitted_order.append(k)
            st.peak_slots = max(st.peak_slots, used)
            st.peak_cases = max(st.peak_cases, len(active))
            if not active:
                sleep(max(0.0, reqs[pending[0]].arrival - (clock() - t0)))
                continue
            left = b.steps_left()
            n = int(min(left[c] for c in active))
            if n > 0:
                b.pair_total(reset=True)
                b.step(min(n, self.chunk), sync=False)
                st.pairs += b.pair_total(reset=True)
                st.chunks += 1
                left = b.steps_left()
            for cid in [c for c in active if left[c] == 0]:
                k = active.pop(cid)
                host = b.download(cid)
                results[k] = Result(k, host, reqs[k].arrival, admitted_at[k], clock() - t0)
                used -= reqs[k].n_particles
                b.remove(cid)
        st.makespan = clock() - t0
        return results, st


def summary(results, stats: Stats) -> dict:
    """p50 / p99 / mean / max time to completion and the aggregate rate."""
    lat = [r.latency for r in results if r is not None]
    return {"p50": percentile(lat, 0.5), "p99": percentile(lat, 0.99),
            "mean": sum(lat) / len(lat) if lat else None, "max": max(lat) if lat else None,
            "makespan_s": stats.makespan, "pairs": stats.pairs, "pair_interactions_per_s": stats.rate(),
            "chunks": stats.chunks, "peak_slots": stats.peak_slots, "peak_cases": stats.peak_cases}
