"""Benchmark: pair-interactions/s of the 2D dam break on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|small|mix|serve|serve-sweep|3d]
                    [--math exact|fast]

Workload "large" (default): 1.0 x 2.0 m water column in a 4 x 3 m tank,
dx = 0.001 (2,000,000 fluid + 30,018 wall particles), Wendland C2, h = 2dx,
Tait EOS, artificial viscosity, Float32; one step = the dam-break loop body
(continuity + EOS, momentum, gravity + move, cell-list build + neighbour
search + kernel + continuity) on state resident in HBM.

Prints one JSON line.  value = sum over the timed steps of the accepted
directed pairs of each step's search (the reference's sum of nCount) / wall
time of those steps.  roofline: the kernel with the largest share of device
time, its algorithmic bytes per launch / its average HIP-event duration.
cpu_baseline: the C restatement of the reference algorithm (reference layout,
OpenMP over the host's cores) on a bounded sample of the same workload, and
the small dam break's full 1000 steps.  The default line also carries the
64-case uneven mix ("mix": batched against one by one) and the serving stream
("serve": p50 / p99 time to completion beside the aggregate rate).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

# the CPU baseline's OpenMP threads stay on their cores (read when the runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "etherparallelparticles.jl_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle of each kernel family: every per-particle
# array the kernel needs, read once and written once (neighbour re-reads and
# the cell-order move are implementation cost and are not counted), plus the
# cell-start array read once per cell.
ALG_BYTES = {
    # the search with the next step's continuity fused (every step of a call but its last; dDensity
    # is zero there and is neither read nor written): pos_nb 8, vel 8, mass 4, vol 4, rho 4, tag 4,
    # h 4, cell 4 in; count 4 out; rho 4, p 4, vol 4 out -- the 56 B of the step's pass after the
    # move -- plus 4 B/cell of cell table read and 4 B/cell of counts zeroed
    "search": (56, 8),
    # momentum with gravity + move and the cell keys fused (the steps after a call's first: dVelocity
    # is zero and is neither read nor written): pos_nb 8, vel 8, rho 4, p 4, mass 4, tag 4, h 4,
    # cell 4 in; then pos 8, alive 4, mov 4, cell 4 in; pos 8, vel 8, key 4, slot 4 out; 4 B/cell
    # of counts
    "momentum": (84, 4),
    # alive 4, tag 4, rho 4, drho 4, mass 4 in; rho 4, drho 4, vol 4, p 4 out
    "continuity_eos": (36, 0),
    # alive 4, tag 4, dvel 8, pos 8, vel 8 in; dvel 8, pos 8, vel 8 out
    "gravity_move": (56, 0),
}


# The 3D pair kernels (compulsory bytes per particle, per cell): the search with kernel + continuity reads
# x, y, z, tag, u, v, w, rho, H, Mass, Volume, the cell and dDensity and writes dDensity and the count (64 B),
# with the cell starts and cell tags (8 B/cell); momentum reads x, y, z, tag, u, v, w, rho, p, Mass, H, the cell
# and du and writes du (72 B) with the same cell tables.
ALG_BYTES_3D = {"search": (64, 8), "momentum": (72, 8), "continuity_eos": (36, 0), "gravity_move": (76, 0)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_times(par, names):
    from eppx import _lib
    L = _lib.lib()
    out = {}
    for name in names:
        avg, cnt = C.c_double(), C.c_int64()
        _lib.check(L.eppx_kernel_time(par.ctx, name.encode(), C.byref(avg), C.byref(cnt), 0), par.ctx, "time")
        if cnt.value:
            out[name] = (avg.value, cnt.value)
    return out


def cpu_baseline(case, seconds, threads, max_steps=200):
    """Time the reference-algorithm restatement (oracle/): whole steps after one
    warm-up step, until `seconds` of CPU time are spent (at least one step)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from eppx import cases

    full = cases.dam_break(case.dx, case.h_factor, case.kernel, M=64)  # reference layout with M = 64 lists
    h = full.host
    ref = O.System(h.int_properties_, h.float_properties_, h.is_alive_, h.cell_index_, h.n_particles_,
                   O.make_layout(full.int_nt, full.float_nt, h.n_capacity), O.domain2d(*full.domain.args),
                   np.array([[1, 1], [1, 0]]), 64)
    ph = O.physics(**full.physics.as_dict())
    ref.search(O.CRIT_SYMMETRY, 1, ph, serial=False, threads=threads)
    ref.run_steps(1, ph, serial=False, threads=threads)
    total, steps, dt = 0.0, 0, 0.0
    while steps < max_steps and (steps == 0 or dt < seconds):
        t0 = time.perf_counter()
        err, pairs = ref.run_steps(1, ph, serial=False, threads=threads)
        dt += time.perf_counter() - t0
        total += float(pairs.sum())
        steps += 1
    return total / dt, dt, steps


def cpu_info(threads: int) -> dict:
    """Model, logical CPUs and physical cores of the CPUs this process may run on."""
    cpus = sorted(os.sched_getaffinity(0))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores), "threads_used": threads,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES")}


def cpu_small_full(threads: int, steps: int = 1000) -> dict:
    """The small dam break (6,518 particles, cubic spline) run in full on the CPU restatement."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from eppx import cases
    c = cases.small(M=40)
    h = c.host
    ref = O.System(h.int_properties_, h.float_properties_, h.is_alive_, h.cell_index_, h.n_particles_,
                   O.make_layout(c.int_nt, c.float_nt, h.n_capacity), O.domain2d(*c.domain.args),
                   np.array([[1, 1], [1, 0]]), 40)
    ph = O.physics(**c.physics.as_dict())
    ref.search(O.CRIT_SYMMETRY, 1, ph, serial=False, threads=threads)
    t0 = time.perf_counter()
    err, pairs = ref.run_steps(steps, ph, serial=False, threads=threads)
    dt = time.perf_counter() - t0
    return {"steps": steps, "seconds": dt, "value": float(pairs.sum()) / dt, "unit": "pair-interactions/s",
            "steps_per_s": steps / dt, "threads": threads}


def copy_calibration(n_bytes: int, repeats: int = 3) -> None:
    """Device-to-device copies of n_bytes (HIP's own copy kernel): the known-size
    reference that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n_bytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n_bytes))
    for _ in range(repeats):
        if hip.hipMemcpy(dst, src, C.c_size_t(n_bytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n_bytes, "repeats": repeats}), flush=True)


def valu_calibration(blocks: int = 8192, iters: int = 20000) -> dict:
    """The VALU calibration kernel (8 x iters FP32 FMAs per lane, blocks x 4 waves): its measured duration
    against the issue-rate model of one wave64 VALU instruction per 4 cycles per SIMD (1024 SIMDs at the
    2.4 GHz peak clock).  Run under rocprofv3 --pmc, it gives the counter readings of a kernel whose VALU
    instruction count is known."""
    import eppx
    from eppx import _lib
    par = eppx.Environment.Parallel(device=0)
    us = C.c_double()
    _lib.check(_lib.lib().eppx_valu_calibration(par.ctx, blocks, iters, C.byref(us)), par.ctx, "valu_calibration")
    waves = blocks * 4
    valu_per_wave = 8 * iters
    model_us = waves * valu_per_wave * 4 / 1024 / 2.4e9 * 1e6
    out = {"valu_calibration": {"blocks": blocks, "waves": waves, "valu_per_wave": valu_per_wave,
                                "launch_us": us.value, "model_us_at_2.4GHz": model_us,
                                "issue_fraction_of_model": model_us / us.value}}
    print(json.dumps(out), flush=True)
    par.close()
    return out


def valu_issue_rate(par, blocks: int = 8192, iters: int = 2000) -> float:
    """Wave-level VALU instructions per second the calibration kernel reaches on this GPU now (the 100 % mark of
    roofline.valu_busy): 8 x iters FP32 FMAs per lane over blocks x 4 waves, timed with HIP events; the median
    of five launches, since the clock the chip holds moves by some 10 % between launches."""
    from eppx import _lib
    times = []
    for _ in range(5):
        us = C.c_double()
        _lib.check(_lib.lib().eppx_valu_calibration(par.ctx, blocks, iters, C.byref(us)), par.ctx,
                   "valu_calibration")
        times.append(us.value)
    med = sorted(times)[len(times) // 2]
    log("VALU calibration kernel (us): " + ", ".join(f"{t:.1f}" for t in times))
    return blocks * 4 * (8 * iters + 33) / (med * 1e-6)


def profile_record(kernel: str, workload: str, math: str):
    """HBM bytes and VALU instructions per launch of a kernel family, from the rocprofv3 counter runs committed
    under profiles/ (counters cannot be read inside this run); the record names the commit it was taken at."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}_{workload}_{math}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _warm_up(par):
    """Loads the kernels and the graph machinery once (a tiny batch and a tiny case)."""
    from eppx import Algorithm, Batch, Class, SPH, cases
    c = cases.dam_break(0.04, 2.0, SPH.WendlandC2, M=1, fluid_width=0.4, fluid_height=0.4)
    b = Batch.Batch(par, c.host, 2 * c.n_particles, 2, cases.ACTIVE_PAIR, max_neighbour_number=64)
    b.add(c.domain, c.physics, c.host, 3)
    b.step(4)
    b.close()
    ps = Class.ParticleSystem(par, c.domain, c.n_particles, c.int_nt, c.float_nt, c.basic_parameters)
    Class.syncto(ps, c.host)
    ns = Class.NeighbourSystem(par, c.domain, cases.ACTIVE_PAIR, max_neighbour_number=64)
    Algorithm.search(ps, c.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(c.physics))
    Algorithm.step(ps, ns, c.physics, n_steps=3)
    ps.close()
    ns.close()


def run_mix(args, par):
    """64 uneven dam-break cases (seed 0), mix_steps steps each: one batch (every case in one segmented cell
    list, one launch per kernel per step, recorded graph) against the same cases one by one on the
    particle-system path.  Both timings cover upload, first search, the steps and download.  Returns the
    measurement."""
    import eppx
    from eppx import Algorithm, Batch, Class, SPH, Serve, cases
    cs = cases.mix_cases(0, steps=args.mix_steps, M=1)
    total_n = sum(c.n_particles for c in cs)
    log(f"mix: {len(cs)} cases, {total_n} particles ({min(c.n_particles for c in cs)}..."
        f"{max(c.n_particles for c in cs)}), {args.mix_steps} steps each")
    par.synchronize()
    t0 = time.perf_counter()
    b = Batch.Batch(par, cs[0].host, total_n, len(cs), cases.ACTIVE_PAIR, max_neighbour_number=64)
    ids = [b.add(c.domain, c.physics, c.host, c.steps) for c in cs]
    ta = time.perf_counter()
    tune = b.autotune(args.tune_trials) if args.tune_trials else None
    tb = time.perf_counter()
    b.pair_total(reset=True)
    par.synchronize()
    t1 = time.perf_counter()
    b.step(args.mix_steps)
    t2 = time.perf_counter()
    pairs_batch = b.pair_total(reset=True)
    if (b.steps_left()[ids] != 0).any():
        raise RuntimeError("a case of the batch did not finish")
    for i in ids:
        b.download(i)
    t3 = time.perf_counter()
    b.close()
    log(f"batch: {t3 - t0:.3f} s end to end, steps {t2 - t1:.3f} s, {pairs_batch} pairs")

    L = eppx._lib.lib()
    done, pairs_single, steps_single = [], 0, 0.0
    t0s = time.perf_counter()
    for c in cs:
        ps = Class.ParticleSystem(par, c.domain, c.n_particles, c.int_nt, c.float_nt, c.basic_parameters)
        Class.syncto(ps, c.host)
        ns = Class.NeighbourSystem(par, c.domain, cases.ACTIVE_PAIR, max_neighbour_number=64)
        Algorithm.search(ps, c.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                         action=SPH.sComputeKernelAndiContinuity(c.physics))
        total = C.c_int64()
        eppx._lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), par.ctx, "pair_total")
        par.synchronize()
        ts = time.perf_counter()
        Algorithm.step(ps, ns, c.physics, n_steps=c.steps)
        steps_single += time.perf_counter() - ts
        eppx._lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), par.ctx, "pair_total")
        pairs_single += total.value
        host = Class.mirror(ps)
        Class.syncto(host, ps)
        ps.close()
        ns.close()
        done.append(time.perf_counter() - t0s)
    t_single = time.perf_counter() - t0s
    log(f"one by one: {t_single:.3f} s end to end, steps {steps_single:.3f} s, {pairs_single} pairs")
    if pairs_single != pairs_batch:
        raise RuntimeError(f"pair counts differ: batch {pairs_batch}, one by one {pairs_single}")
    t_batch = (t3 - t0) - (tb - ta)  # the autotuner's trials are reported apart
    return {
        "metric": "aggregate pair-interactions/s, 64-case uneven 2D dam-break mix, batched",
        "value": pairs_batch / t_batch, "unit": "pair-interactions/s", "n_gpus": 1, "steps": args.mix_steps,
        "warmup": 0, "ms_per_step": 1000.0 * (t2 - t1) / args.mix_steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"64 dam-break cases from seed 0 (dx in {{0.02, 0.01, 0.005, 0.0025}}, width "
                               f"U[0.5, 1.5], aspect U[0.5, 2], Wendland C2, h = 2dx), {args.mix_steps} steps each",
                   "particles": total_n, "cases": len(cs), "math": args.math, "seed": 0,
                   "case_particles_min": min(c.n_particles for c in cs),
                   "case_particles_max": max(c.n_particles for c in cs)},
        "one_by_one": {"value": pairs_single / t_single, "seconds": t_single, "steps_seconds": steps_single},
        "batched": {"seconds": t_batch, "steps_seconds": t2 - t1, "add_seconds": (ta - t0),
                    "download_seconds": t3 - t2, "autotune_seconds": tb - ta,
                    "pair_shape": {"search": tune["search"], "momentum": tune["momentum"]} if tune else None},
        "speedup_end_to_end": t_single / t_batch,
        "speedup_steps": steps_single / (t2 - t1),
        "latency_s": {"batched_p50": t_batch, "batched_p99": t_batch,
                      "one_by_one_p50": Serve.percentile(done, 0.5), "one_by_one_p99": Serve.percentile(done, 0.99)},
        "pairs": pairs_batch, "pairs_one_by_one": pairs_single, "pairs_equal": pairs_single == pairs_batch,
    }


def run_serve(args, par, rate):
    """Serving stream (seed 1): Poisson arrivals at `rate` per second of dam-break requests (N log-uniform in
    [1e4, 1e6], dx from a 1 x 2 m column, steps U{200..2000}), served by eppx.Serve (continuous batching:
    admission by device bytes per particle against a fixed HBM budget, cases join between chunks and leave
    when done).  Returns p50 / p99 time to completion and the aggregate pair-interactions/s."""
    from eppx import Serve, cases
    reqs = cases.serving_requests(1, n_requests=args.serve_requests, rate_hz=rate)
    budget = int(args.serve_budget_gb * 1e9)
    log(f"serve: {len(reqs)} requests at {rate}/s over {reqs[-1].arrival:.2f} s, "
        f"{sum(r.n_particles for r in reqs)} particles, budget {args.serve_budget_gb} GB")
    srv = Serve.Server(par, reqs[0].host, budget, args.serve_slots, cases.ACTIVE_PAIR, max_neighbour_number=64,
                       chunk=args.serve_chunk)
    try:
        par.synchronize()
        results, stats = srv.run(reqs)
    finally:
        srv.close()
    s = Serve.summary(results, stats)
    log(f"serve at {rate}/s: p50 {s['p50']:.3f} s, p99 {s['p99']:.3f} s, makespan {s['makespan_s']:.2f} s, "
        f"{s['pair_interactions_per_s'] / 1e9:.1f} G pair-interactions/s")
    return {
        "metric": "serving time-to-completion p50/p99, Poisson 2D dam-break stream (continuous batching)",
        "value": s["pair_interactions_per_s"], "unit": "pair-interactions/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(reqs)} requests, seed 1, Poisson at {rate}/s, N log-uniform in [1e4, 1e6] "
                               f"(dx = sqrt(2/N)), steps U{{200..2000}}, Wendland C2, h = 2dx",
                   "seed": 1, "rate_per_s": rate, "math": args.math,
                   "particles": sum(r.n_particles for r in reqs), "steps_total": sum(r.steps for r in reqs)},
        "latency_s": {"p50": s["p50"], "p99": s["p99"], "mean": s["mean"], "max": s["max"]},
        "admission": {"bytes_per_particle": srv.bytes_per_slot, "budget_bytes": srv.slots * srv.bytes_per_slot,
                      "slots": srv.slots, "case_slots": args.serve_slots, "peak_slots": s["peak_slots"],
                      "peak_cases": s["peak_cases"]},
        "makespan_s": s["makespan_s"], "chunks": s["chunks"], "pairs": s["pairs"],
    }


def cpu_baseline_3d(seconds, threads, dx=0.015, max_steps=50):
    """The 3D restatement (oracle/eppx_oracle3d.c) on the 3D dam break's geometry at dx (same h / dx, so the
    same ~256 neighbours per particle as the benchmark case): whole steps after one search, until `seconds`."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import oracle3d as O3
    from eppx import cases
    c = cases.dam_break_3d(dx, M=300)
    h = c.host
    ref = O3.System3D(h.int_properties_, h.float_properties_, h.is_alive_, h.cell_index_, h.n_particles_,
                      O.make_layout(c.int_nt, c.float_nt, h.n_capacity), O3.domain3d(*c.domain.args),
                      np.array([[1, 1], [1, 0]]), 300)
    p = c.physics
    ph = O3.physics3d(gz=p.gz, **{k: v for k, v in p.as_dict().items() if k != "gz"})
    ref.search(O.CRIT_SYMMETRY, 1, ph, serial=False, threads=threads)
    total, steps, dt = 0.0, 0, 0.0
    while steps < max_steps and (steps == 0 or dt < seconds):
        t0 = time.perf_counter()
        err, pairs = ref.run_steps(1, ph, serial=False, threads=threads)
        dt += time.perf_counter() - t0
        total += float(pairs.sum())
        steps += 1
    return total / dt, dt, steps, c.n_particles


def run_3d(args, par):
    """The 3D dam break (1.0 x 0.5 x 1.0 m block in a 3.2 x 1.0 x 1.8 m tank, dx = 0.0075: 2,180,764 particles,
    Wendland C2 3D, h = 2dx, Tait EOS, artificial viscosity): args.steps steps (100 by default) of the loop body
    on state resident in HBM, pair-interactions/s, the dominant kernel's roofline and the 3D restatement on the
    host cores."""
    from eppx import Algorithm, SPH, _lib, cases
    L = _lib.lib()
    steps = args.steps if args.steps_given else 100
    t_setup = time.perf_counter()
    case = cases.large_3d()
    log(f"3D case: {case.n_particles} particles, {case.domain.get_n()} cells ({time.perf_counter() - t_setup:.1f}s)")
    from eppx import Class
    ps = Class.ParticleSystem(par, case.domain, case.n_particles, case.int_nt, case.float_nt, case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(par, case.domain, cases.ACTIVE_PAIR, max_neighbour_number=400)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    Algorithm.step(ps, ns, case.physics, n_steps=args.warmup)
    log(f"warmup {args.warmup} steps done")
    total = C.c_int64()
    _lib.check(L.eppx_pair_total3d(ps.handle, C.byref(total), 1), par.ctx, "pair_total3d")
    par.synchronize()
    t0 = time.perf_counter()
    Algorithm.step(ps, ns, case.physics, n_steps=steps, sync=False)
    par.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.check(L.eppx_pair_total3d(ps.handle, C.byref(total), 1), par.ctx, "pair_total3d")
    pairs = total.value
    alive = ps.get_alive_n_particles()
    _lib.check(L.eppx_set_timing(par.ctx, 1), par.ctx, "timing")
    L.eppx_kernel_time(par.ctx, b"search", None, None, 1)
    Algorithm.step(ps, ns, case.physics, n_steps=args.profile_steps_3d)
    _lib.check(L.eppx_set_timing(par.ctx, 0), par.ctx, "timing")
    names = ["search", "momentum", "continuity_eos", "gravity_move", "cell_key", "cell_scan", "cell_permute"]
    kt = kernel_times(par, names)
    log("kernel averages (us): " + ", ".join(f"{k}={v[0]:.1f}x{v[1]}" for k, v in kt.items()))
    dominant = max((k for k in kt if k in ALG_BYTES_3D), key=lambda k: kt[k][0] * kt[k][1])
    per_p, per_cell = ALG_BYTES_3D[dominant]
    alg_bytes = per_p * case.n_particles + per_cell * case.domain.get_n()
    avg_us = kt[dominant][0]
    achieved = alg_bytes / (avg_us * 1e-6) / 1e9
    device_us_per_step = sum(v[0] * v[1] for v in kt.values()) / args.profile_steps_3d
    prof = profile_record(dominant, "3d", args.math)
    issue_rate = valu_issue_rate(par)
    valu = prof["valu_insts_per_launch"] / (avg_us * 1e-6 * issue_rate) if prof.get("valu_insts_per_launch") else None
    frac = achieved / HBM_PEAK_GBS
    limiter = None if valu is None else ("valu_issue" if valu >= 0.6 else ("hbm" if frac >= 0.6 else "latency"))
    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        host = cpu_info(threads)  # before OpenMP binds this thread to one core
        rate, secs, nsteps, n_cpu = cpu_baseline_3d(args.cpu_seconds, threads)
        cpu = {"value": rate, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
               "sample": f"{nsteps} steps of the same 3D dam break geometry at dx = 0.015 ({n_cpu} particles, the "
                         f"same h / dx and neighbours per particle) after one search, C/OpenMP 3D restatement on "
                         f"the reference layout (per-cell atomic insert, materialised M = 300 lists), {secs:.1f} s",
               "host": host}
        log(f"cpu 3D baseline: {rate / 1e9:.3f} G/s on {threads} threads")
    ps.close()
    ns.close()
    ms = 1000.0 * elapsed / steps
    return {
        "metric": "pair-interactions/s at 2.18M-particle 3D dam break",
        "value": pairs / elapsed, "unit": "pair-interactions/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms, "steps_per_s": 1000.0 / ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D dam break: 1.0 x 0.5 x 1.0 m block in a 3.2 x 1.0 x 1.8 m tank, dx = 0.0075, "
                               f"{case.n_particles} particles (3 wall layers), Wendland C2 3D, h = 2dx, Tait EOS, "
                               f"artificial viscosity",
                   "particles": case.n_particles, "alive_after": alive, "cells": case.domain.get_n(),
                   "math": args.math, "pairs_per_step": pairs / steps,
                   "parity": "3D restatement (the reference's 3D path is a stub: unpinned by reference)"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": frac, "traffic": prof.get("hbm_bytes_per_launch"), "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_us": avg_us, "valu_busy": valu, "valu_issue_rate": issue_rate, "limiter": limiter,
                     "counters_commit": prof.get("commit")},
        "device_us_per_step": device_us_per_step,
        "kernels_us": {k: round(v[0], 2) for k, v in kt.items()},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (200; the 3D workload: 100)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["large", "small", "mix", "serve", "serve-sweep", "3d"], default="large")
    ap.add_argument("--profile-steps-3d", type=int, default=5, help="3D: eager steps timed kernel by kernel")
    ap.add_argument("--mix-steps", type=int, default=500)
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=32.0, help="Poisson arrival rate (1/s)")
    ap.add_argument("--serve-sweep", default="4,8,16,32,64", help="arrival rates of --workload serve-sweep")
    ap.add_argument("--no-mix", action="store_true", help="default run: leave out the 64-case mix")
    ap.add_argument("--no-serve", action="store_true", help="default run: leave out the serving stream")
    ap.add_argument("--serve-budget-gb", type=float, default=2.0, help="HBM admitted to resident cases")
    ap.add_argument("--serve-slots", type=int, default=32, help="cases resident at once")
    ap.add_argument("--serve-chunk", type=int, default=100, help="largest number of steps between polls")
    ap.add_argument("--math", choices=["exact", "fast"], default="fast")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length")
    ap.add_argument("--cpu-seconds-1t", type=float, default=5.0,
                    help="single-thread CPU baseline sample length (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=50, help="eager steps timed kernel by kernel")
    ap.add_argument("--tune-trials", type=int, default=3, help="autotuner trials per block size (0: no tuning)")
    ap.add_argument("--pair-shape", default=None, metavar="SB,SS,MB,MS",
                    help="pin the pair-kernel shapes instead of tuning: search block, staged, momentum block, staged")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--valu-calibration", action="store_true", help="only run the VALU calibration kernel")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of BYTES (profiler calibration)")
    args = ap.parse_args()
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 200
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.valu_calibration:
        valu_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "libeppx runs one GPU per process; multi-GPU runs are not part of this "
                                   "build (see DESIGN.md: how the path would shard)", "n_gpus": args.gpus}))
        sys.exit(2)

    import eppx
    from eppx import Algorithm, Class, SPH, _lib, cases

    if not os.path.exists(eppx.library_path()):
        eppx.build()
    if args.workload in ("mix", "serve", "serve-sweep", "3d"):
        par = eppx.Environment.Parallel(device=0, math=args.math)
        _warm_up(par)
        if args.workload == "3d":
            out = run_3d(args, par)
        elif args.workload == "mix":
            out = run_mix(args, par)
        elif args.workload == "serve":
            out = run_serve(args, par, args.serve_rate)
        else:
            sweep = [run_serve(args, par, float(r)) for r in args.serve_sweep.split(",")]
            out = dict(sweep[-1])
            out["metric"] = "serving time-to-completion p50/p99 against arrival rate, Poisson 2D dam-break stream"
            out["sweep"] = [{"rate_per_s": x["config"]["rate_per_s"], "p50_s": x["latency_s"]["p50"],
                             "p99_s": x["latency_s"]["p99"], "pair_interactions_per_s": x["value"],
                             "makespan_s": x["makespan_s"], "peak_cases": x["admission"]["peak_cases"]}
                            for x in sweep]
        print(json.dumps(out), flush=True)
        par.close()
        return
    t_setup = time.perf_counter()
    if args.workload == "large":
        case = cases.dam_break(0.001, 2.0, SPH.WendlandC2, M=1)   # slim host tables; device never stores lists
        M = 64
    else:
        case = cases.dam_break(0.02, 1.5, SPH.CubicSpline, M=1)
        M = 40
    log(f"case: {case.n_particles} particles, {case.domain.get_n()} cells ({time.perf_counter() - t_setup:.1f}s)")
    par = eppx.Environment.Parallel(device=0, math=args.math)
    ps = Class.ParticleSystem(par, case.domain, case.n_particles, case.int_nt, case.float_nt, case.basic_parameters)
    Class.syncto(ps, case.host)
    ns = Class.NeighbourSystem(par, case.domain, cases.ACTIVE_PAIR, max_neighbour_number=M)
    Algorithm.search(ps, case.domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(case.physics))
    L = _lib.lib()
    Algorithm.step(ps, ns, case.physics, n_steps=args.warmup)
    log(f"warmup {args.warmup} steps done")
    if args.pair_shape:
        sb, ss, mb, ms = (int(v) for v in args.pair_shape.split(","))
        Algorithm.set_pair_shape(ps, search=(sb, ss), momentum=(mb, ms))
        tune = None
    else:
        tune = Algorithm.autotune(ps, ns, case.physics, trials=args.tune_trials) if args.tune_trials else None
    if tune:
        log(f"autotune: search {tune['search']}, momentum {tune['momentum']} "
            f"(us: {tune['search_us']}, {tune['momentum_us']})")
    total = C.c_int64()
    _lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), par.ctx, "pair_total")
    fb = C.c_int64()
    _lib.check(L.eppx_pair_fallback_blocks(ps.handle, C.byref(fb), 1), par.ctx, "fallback")

    # timed region: whole steps replayed as a recorded hipGraph, state resident in HBM
    par.synchronize()
    t0 = time.perf_counter()
    Algorithm.step(ps, ns, case.physics, n_steps=args.steps, step_offset=args.warmup, sync=False)
    par.synchronize()
    elapsed = time.perf_counter() - t0

    _lib.check(L.eppx_pair_total(ps.handle, C.byref(total), 1), par.ctx, "pair_total")
    _lib.check(L.eppx_check_finite(ps.handle), par.ctx, "finite")
    pairs = total.value
    _lib.check(L.eppx_pair_fallback_blocks(ps.handle, C.byref(fb), 1), par.ctx, "fallback")
    blocks_per_launch = -(-case.n_particles // 256)  # 256 particles per pair block
    log(f"pair blocks reading HBM instead of LDS: {fb.value / (2 * args.steps) / blocks_per_launch:.4%}")

    # per-kernel durations: the same steps launched kernel by kernel between HIP events on the library's stream
    _lib.check(L.eppx_set_timing(par.ctx, 1), par.ctx, "timing")
    L.eppx_kernel_time(par.ctx, b"search", None, None, 1)  # clear
    Algorithm.step(ps, ns, case.physics, n_steps=args.profile_steps, step_offset=args.warmup + args.steps)
    _lib.check(L.eppx_set_timing(par.ctx, 0), par.ctx, "timing")
    names = ["search", "momentum", "continuity_eos", "gravity_move", "cell_reset", "cell_key", "cell_scan",
             "cell_members", "cell_permute"]
    kt = kernel_times(par, names)
    log("kernel averages (us): " + ", ".join(f"{k}={v[0]:.1f}x{v[1]}" for k, v in kt.items()))
    dominant = max((k for k in kt if k in ALG_BYTES), key=lambda k: kt[k][0] * kt[k][1])
    per_p, per_cell = ALG_BYTES[dominant]
    alg_bytes = per_p * case.n_particles + per_cell * case.domain.get_n()
    avg_us = kt[dominant][0]
    achieved = alg_bytes / (avg_us * 1e-6) / 1e9
    device_us_per_step = sum(v[0] * v[1] for v in kt.values()) / args.profile_steps
    # the whole step's compulsory traffic: 56 B/particle for the pass after the move (x, u, m, tag, h,
    # rho, V in; rho, p, V out; cell key written and read) and 52 B/particle for the momentum pass
    # (x, u, rho, p, m, tag, h in; x, u out), plus the 12 B/cell of the cell table
    step_bytes = 108 * case.n_particles + 12 * case.domain.get_n()

    # HBM bytes per launch of the dominant kernel and its VALU instructions per launch, from the rocprofv3
    # counter runs committed under profiles/; VALU busy = those instructions over this run's launch duration
    # times the issue rate the calibration kernel reaches now
    prof = profile_record(dominant, args.workload, args.math)
    traffic, prof_commit = prof.get("hbm_bytes_per_launch"), prof.get("commit")
    issue_rate = valu_issue_rate(par)
    valu = prof["valu_insts_per_launch"] / (avg_us * 1e-6 * issue_rate) if prof.get("valu_insts_per_launch") else None
    frac = achieved / HBM_PEAK_GBS
    limiter = None
    if valu is not None:
        limiter = "valu_issue" if valu >= 0.6 else ("hbm" if frac >= 0.6 else "latency")

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        host = cpu_info(threads)  # before OpenMP binds this thread to one core
        rate, secs, nsteps = cpu_baseline(case, args.cpu_seconds, threads)
        cpu = {"value": rate, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
               "sample": f"{nsteps} steps of the same {case.n_particles}-particle dam break after one "
                         f"warm-up step, C/OpenMP restatement of the reference algorithm on the reference layout "
                         f"(per-cell atomic insert, materialised M=64 neighbour lists), {secs:.1f} s"}
        cpu["host"] = host
        if args.cpu_seconds_1t > 0 and threads > 1:
            rate1, secs1, n1 = cpu_baseline(case, args.cpu_seconds_1t, 1)
            cpu["single_thread"] = {"value": rate1, "cores": 1,
                                    "sample": f"{n1} steps after one warm-up step on one thread, {secs1:.1f} s"}
            log(f"cpu baseline: {rate / 1e9:.3f} G/s on {threads} threads, {rate1 / 1e9:.3f} G/s on 1")
        cpu["small_dam_break_1000_steps"] = cpu_small_full(threads)
        log(f"cpu small dam break, 1000 steps: {cpu['small_dam_break_1000_steps']['seconds']:.2f} s")

    ms = 1000.0 * elapsed / args.steps
    out = {
        "metric": "pair-interactions/s at 2M-particle 2D dam break",
        "value": pairs / elapsed,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "steps_per_s": 1000.0 / ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"2D dam break {args.workload}: {case.n_particles} particles (dx={case.dx}, "
                               f"{'Wendland C2' if args.workload == 'large' else 'cubic spline'}, "
                               f"h={case.h_factor}dx, Tait EOS, artificial viscosity)",
                   "particles": case.n_particles, "cells": case.domain.get_n(), "math": args.math,
                   "pairs_per_step": pairs / args.steps,
                   "pair_shape": {"search": tune["search"], "momentum": tune["momentum"]} if tune else None,
                   "untiled_block_fraction": fb.value / (2 * args.steps) / blocks_per_launch},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": frac, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_us, "valu_busy": valu,
                     "valu_issue_rate": issue_rate, "limiter": limiter, "counters_commit": prof_commit},
        "step_roofline": {"alg_bytes_per_step": step_bytes, "device_us_per_step": device_us_per_step,
                          "hbm_frac": step_bytes / (device_us_per_step * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "kernels_us": {k: round(v[0], 2) for k, v in kt.items()},
        "cpu_baseline": cpu,
    }
    if args.workload == "large" and not args.no_mix:
        m = run_mix(args, par)
        out["mix"] = {"value": m["value"], "unit": m["unit"], "one_by_one_value": m["one_by_one"]["value"],
                      "speedup_end_to_end": m["speedup_end_to_end"], "speedup_steps": m["speedup_steps"],
                      "batched_seconds": m["batched"]["seconds"], "one_by_one_seconds": m["one_by_one"]["seconds"],
                      "ms_per_batch_step": m["ms_per_step"], "pairs": m["pairs"], "pairs_equal": m["pairs_equal"],
                      "config": m["config"]}
    if args.workload == "large" and not args.no_serve:
        v = run_serve(args, par, args.serve_rate)
        out["serve"] = {"value": v["value"], "unit": v["unit"], "latency_s": v["latency_s"],
                        "makespan_s": v["makespan_s"], "admission": v["admission"], "config": v["config"]}
    print(json.dumps(out), flush=True)
    par.close()


if __name__ == "__main__":
    main()
