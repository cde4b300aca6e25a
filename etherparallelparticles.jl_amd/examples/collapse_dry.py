"""The collapse_dry/same example on eppx: casegen.jl writes a case as a
configuration plus a first raw file, main.jl loads them and runs the
dam-break loop with the Shepard filter every 20 steps and an output every
400 (example/collapse_dry/same/casegen.jl, main.jl:148-207).

    python etherparallelparticles.jl_amd/examples/collapse_dry.py casegen OUT [--dx 0.02]
    python etherparallelparticles.jl_amd/examples/collapse_dry.py main OUT [--steps N]

The physics constants live in the driver, as in main.jl (they are not part of
the configuration).  --physics benchmark (the default): Tait EOS, artificial
viscosity, dt = 0.1 h / c0, g = 9.81.  --physics example: main.jl's own
constants (main.jl:57-70): linear EOS with c0 = 120, classic viscosity with
mu0 = 1e-3, g = 9.8, dt = 0.2 h0 / c0.  Both use the kernel and h of the case.
"""
from __future__ import annotations

import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from eppx import Algorithm, Class, DataIO, SPH, cases  # noqa: E402

KERNELS = {"cubic": SPH.CubicSpline, "wendland": SPH.WendlandC2}


def driver_physics(drv: dict) -> SPH.Physics:
    """The step's constants: main.jl's own (physics "example") or the benchmark's (the default)."""
    if drv.get("physics", "benchmark") == "example":
        return SPH.example_physics(drv["dx"], drv["h_factor"], KERNELS[drv["kernel"]])
    return SPH.dam_break_physics(drv["dx"], drv["h_factor"], KERNELS[drv["kernel"]])


def casegen(out: str, dx: float = 0.02, h_factor: float = 1.5, kernel: str = "cubic", M: int = 40,
            physics: str = "benchmark") -> str:
    """casegen.jl: the dam break as a configuration (config/config.json) and
    the raw file of write step 0.  Returns the configuration's path."""
    case = cases.dam_break(dx, h_factor, KERNELS[kernel], M=M)
    par = DataIO.Parallel(DataIO.template())
    writer = DataIO.Writer(out)
    ns = type("NeighbourSystemRecord", (), {"active_pair_": Class.ActivePair(cases.ACTIVE_PAIR),
                                             "max_neighbour_number": M,
                                             # casegen.jl:151
                                             "periodic_boundary_": Class.PeriodicBoundary(
                                                 par, case.domain, Class.NonePeriodicBoundaryPolicy)})()
    config = DataIO.config(par, case.domain, case.host, ns, writer)
    # the physics main.jl hard-codes
    config["driver"] = {"dx": dx, "h_factor": h_factor, "kernel": kernel, "physics": physics}
    DataIO.save(writer, case.host, DataIO.appendix())
    return DataIO.writeConfig(writer, config)


def main(out: str, total_steps: int, filter_interval: int = 20, output_interval: int = 400, step: int = 0,
         math: str = "exact", log=print) -> dict:
    """main.jl: load the configuration and the raw file of write step `step`,
    then run total_steps steps, filtering every filter_interval and writing
    the state every output_interval (counted from the start of the case).
    Returns the last appendix."""
    config = DataIO.parse(os.path.join(out, "config", "config.json"))
    drv = config["driver"]
    par = DataIO.Parallel(config, math=math)
    domain = DataIO.Domain(config, par)
    writer = DataIO.WriterFrom(config)
    ps = DataIO.ParticleSystem(config, par, domain)  # n_particles rows, filled below
    host = Class.mirror(ps)
    appendix = DataIO.load_into(writer, host, step)
    Class.syncto(ps, host)
    ns = DataIO.NeighbourSystem(config, par, domain)
    phys = driver_physics(drv)
    t, tm, write_step = float(appendix["TimeValue"]), int(appendix["TMSTEP"]), int(appendix["WriteStep"])
    Algorithm.search(ps, domain, ns, criterion=Algorithm.symmetrySearchCriterion,
                     action=SPH.sComputeKernelAndiContinuity(phys))
    end = tm + total_steps
    while tm < end:
        # up to the next output step, in one call (filter steps inside it)
        n = min(end, (tm // output_interval + 1) * output_interval) - tm
        Algorithm.step(ps, ns, phys, n_steps=n, filter_interval=filter_interval, step_offset=tm)
        tm += n
        t += n * phys.dt
        if tm % output_interval == 0:
            write_step += 1
            appendix.update(TimeValue=t, TMSTEP=tm, WriteStep=write_step)
            name = DataIO.syncsave(writer, host, ps, appendix)
            log(f"step {tm}: t = {t:.6f} s, {ps.get_alive_n_particles()} alive -> {name}")
    return appendix


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("what", choices=["casegen", "main"])
    ap.add_argument("out")
    ap.add_argument("--dx", type=float, default=0.02)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--math", choices=["exact", "fast"], default="exact")
    ap.add_argument("--physics", choices=["benchmark", "example"], default="benchmark",
                    help="casegen: the step's constants (the benchmark's, or main.jl's own)")
    a = ap.parse_args()
    if a.what == "casegen":
        print(casegen(a.out, dx=a.dx, physics=a.physics))
    else:
        main(a.out, a.steps, math=a.math)
