"""DataIO: the reference's JSON configuration and raw output, mapped onto eppx.

Configuration (src/DataIO/Config/Input.jl:10-114, Output.jl:11-186, the
section and key names of template.json:1-83): a dict with the sections
"parallel", "domain", "particle_system", "neighbour_system" and "writer".
``Parallel``, ``Domain``, ``ParticleSystem``, ``NeighbourSystem`` and
``Writer`` build handles from it (Input.jl); ``config`` writes handles back
into it and ``writeConfig`` stores it (Output.jl).

Nothing in a configuration is evaluated.  The reference's
``capacity_expand`` field holds Julia source (``"n -> n + 100"``, written by
Output.jl:57); ``capacity_expand`` parses it as an integer expression in n
(literals, n, + - * and integer division ÷ / div, parentheses) and refuses
anything else.

Raw output (src/DataIO/Writer/RawWriter.jl:10-56, Writer.jl:156-241): the
reference writes JLD2 (an HDF5 profile) with the datasets
``raw/int_properties``, ``raw/float_properties`` (alive rows only) and
``appendix/<key>``.  There is no HDF5 library in this build, so the same
datasets go into a NumPy ``.npz`` archive under the same names; a writer
configured with the ``.jld2`` suffix writes ``.npz`` instead (with a
warning), and loading a ``.jld2`` file raises.
"""
from __future__ import annotations

import ast
import copy
import datetime
import json
import os
import re
import warnings
from collections import OrderedDict

import numpy as np

from . import Class, Environment

# Backend names of Environment.kNameToBackend (src/Environment/Environment.jl);
# every one of them maps to the HIP device behind libeppx.
BACKENDS = ("cpu", "cuda", "rocm", "oneapi", "metal")


def template() -> OrderedDict:
    """A configuration with every section and key of the reference template
    (template.json:1-83), holding the 2D dam-break layout of
    example/collapse_dry/same/casegen.jl with M = 50 neighbour slots."""
    m = 50
    return OrderedDict([
        ("parallel", OrderedDict([("int", "Int32"), ("float", "Float32"), ("backend", "rocm"), ("device", 1)])),
        ("domain", OrderedDict([("dimension", 2), ("gap", 0.03), ("first_x", -0.03), ("first_y", -0.03),
                                ("first_z", 0.0), ("last_x", 4.03), ("last_y", 3.03), ("last_z", 0.0)])),
        ("particle_system", OrderedDict([
            ("n_particles", 0),
            ("capacity_expand", "n -> n"),
            ("load", OrderedDict([("mode", "initial"), ("file", "")])),
            ("int_named_tuple", OrderedDict([("Tag", 1), ("IsMovable", 1), ("nCount", 1), ("nIndex", m)])),
            ("float_named_tuple", OrderedDict([
                ("PositionVec", 2), ("VelocityVec", 2), ("dVelocityVec", 2), ("AccelerationVec", 2), ("Mass", 1),
                ("Volume", 1), ("Density", 1), ("dDensity", 1), ("Pressure", 1), ("Gap", 1), ("H", 1),
                ("SumWeight", 1), ("SumWeightedDensity", 1), ("SumWeightedPressure", 1), ("nW", m), ("nDW", m),
                ("nHInv", m), ("nRVec", 2 * m), ("nR", m)])),
            ("basic_parameters", OrderedDict([("rho0", 1e3), ("p0", 0.0), ("FLUID_TAG", 1), ("WALL_TAG", 2),
                                              ("gap0", 0.01), ("h0", 0.015), ("mu0", 1e-3)])),
            ("basic_index_map_dict", OrderedDict()),
        ])),
        ("neighbour_system", OrderedDict([
            ("active_pair", [[1, 1], [1, 2], [2, 1]]),
            ("periodic_boundary_policy", OrderedDict([("type", "none"), ("axis", [False, False, False])])),
            ("max_neighbour_number", m),
            ("n_threads", 256),
        ])),
        ("writer", OrderedDict([("path", "result"), ("file_name", "result"), ("connect", "_"), ("digits", 4),
                                ("suffix", ".npz")])),
    ])


def parse(path: str) -> OrderedDict:
    """JSON.parsefile(path; dicttype = OrderedDict), as main.jl:19 reads a configuration."""
    with open(path) as f:
        return json.load(f, object_pairs_hook=OrderedDict)


# ------------------------------------------------------------------ Input.jl
def Parallel(config: dict, math: str = "exact") -> Environment.Parallel:
    """DataIO.Parallel (Input.jl:10-19).  IT / FT must be Int32 / Float32; the
    backend name may be any of the reference's (the device is always the HIP
    device behind libeppx); "device" is the reference's 1-based device number."""
    p = config["parallel"]
    if p["int"] != "Int32" or p["float"] != "Float32":
        raise NotImplementedError(f"libeppx computes in Int32 / Float32, not {p['int']} / {p['float']}")
    if p["backend"] not in BACKENDS:
        raise KeyError(f"unknown backend {p['backend']!r} (one of {', '.join(BACKENDS)})")
    device = int(p.get("device", 1))
    if device < 1:
        raise ValueError(f"device numbers start at 1, got {device}")
    return Environment.Parallel(np.int32, np.float32, device=device - 1, math=math)


def Domain(config: dict, parallel: Environment.Parallel | None = None) -> Class.Domain2D:
    """DataIO.Domain (Input.jl:21-37): Domain2D(gap, first_x, first_y, last_x, last_y)."""
    d = config["domain"]
    if int(d["dimension"]) != 2:
        raise NotImplementedError("3D domains are unfinished in the reference (Input.jl:35)")
    return Class.Domain2D(float(d["gap"]), float(d["first_x"]), float(d["first_y"]), float(d["last_x"]),
                          float(d["last_y"]))


_EXPAND = re.compile(r"^\s*n\s*->\s*(.+?)\s*$", re.S)


def capacity_expand(source: str, n: int) -> int:
    """The capacity of n particles under the configuration's capacity_expand
    Julia lambda ("n -> <expr>"), evaluated as integer arithmetic without
    eval: literals, n, + - * ÷ div(a, b), unary minus and parentheses.
    Raises ValueError for anything else."""
    m = _EXPAND.match(source)
    if not m:
        raise ValueError(f"capacity_expand must read 'n -> <expression>', got {source!r}")
    if "//" in m.group(1):
        raise ValueError(f"capacity_expand: '//' is Julia's rational division, not an integer: {source!r}")
    expr = m.group(1).replace("÷", "//")
    expr = re.sub(r"(\d)\s*n\b", r"\1 * n", expr)  # Julia's juxtaposed 2n
    try:
        tree = ast.parse(expr, mode="eval")
    except SyntaxError as e:
        raise ValueError(f"capacity_expand: cannot parse {source!r}") from e

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, int) and not isinstance(node.value, bool):
            return node.value
        if isinstance(node, ast.Name) and node.id == "n":
            return int(n)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.FloorDiv)):
            a, b = ev(node.left), ev(node.right)
            if isinstance(node.op, ast.Add):
                return a + b
            if isinstance(node.op, ast.Sub):
                return a - b
            if isinstance(node.op, ast.Mult):
                return a * b
            if b == 0:
                raise ValueError("capacity_expand: division by zero")
            q = abs(a) // abs(b)  # Julia's ÷ truncates toward zero (integers only: exact at any size)
            return -q if (a < 0) != (b < 0) else q
        if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "div"
                and len(node.args) == 2 and not node.keywords):
            return ev(ast.BinOp(left=node.args[0], op=ast.FloorDiv(), right=node.args[1]))
        raise ValueError(f"capacity_expand: {ast.dump(node)} is not integer arithmetic in n")

    cap = ev(tree)
    if cap < n:
        raise ValueError(f"capacity_expand gives {cap} slots for {n} particles")
    return int(cap)


def _named(d: dict) -> dict:
    return OrderedDict((str(k), int(v)) for k, v in d.items())


def _raw_tables(path: str):
    """(int_properties, float_properties) of a raw file written by save."""
    if path.endswith(".jld2"):
        raise NotImplementedError(f"{path}: JLD2 files are HDF5, and this build has no HDF5 reader; "
                                  "write the state with DataIO.save (an .npz archive of the same datasets)")
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["raw/int_properties"], np.int32), np.asarray(z["raw/float_properties"], np.float32)


def ParticleSystem(config: dict, parallel: Environment.Parallel, domain: Class.Domain2D) -> Class.ParticleSystem:
    """DataIO.ParticleSystem (Input.jl:39-80).  Mode "initial": n_particles
    rows, alive, zero-filled; mode "continue": the rows of the raw file named
    by load.file, uploaded."""
    c = config["particle_system"]
    int_nt, float_nt = _named(c["int_named_tuple"]), _named(c["float_named_tuple"])
    basic = OrderedDict(c["basic_parameters"])
    index_map = OrderedDict(c.get("basic_index_map_dict") or {}) or None
    mode = c["load"]["mode"]
    if mode == "initial":
        n = int(c["n_particles"])
        tables = None
    elif mode == "continue":
        tables = _raw_tables(c["load"]["file"])
        n = tables[0].shape[0]
    else:
        raise ValueError(f"particle_system.load.mode must be 'initial' or 'continue', got {mode!r}")
    cap = capacity_expand(c["capacity_expand"], n)
    ps = Class.ParticleSystem(parallel, domain, n, int_nt, float_nt, basic, n_capacity=cap,
                              basic_index_map_dict=index_map)
    if tables is not None:
        host = Class.mirror(ps)
        set_tables(host, *tables)
        Class.syncto(ps, host)
    return ps


def NeighbourSystem(config: dict, parallel: Environment.Parallel, domain: Class.Domain2D) -> Class.NeighbourSystem:
    """DataIO.NeighbourSystem (Input.jl:82-104)."""
    c = config["neighbour_system"]
    pairs = [(int(a), int(b)) for a, b in c["active_pair"]]
    return Class.NeighbourSystem(parallel, domain, pairs, periodic_policy(c["periodic_boundary_policy"], domain),
                                 max_neighbour_number=int(c["max_neighbour_number"]),
                                 n_threads=int(c.get("n_threads", 256)))


def periodic_policy(section: dict, domain) -> type:
    """The policy of a configuration's `periodic_boundary_policy` section
    (Input.jl:84-92): "none", else the policy of the domain's dimension with
    the first N flags of `axis`."""
    if section["type"] == "none":
        return Class.NonePeriodicBoundaryPolicy
    axis = section["axis"]
    if not all(isinstance(a, bool) for a in axis):
        raise ValueError(f"periodic_boundary_policy.axis must hold booleans, got {axis!r}")
    if domain.dimension() == 2:
        return Class.PeriodicBoundaryPolicy2D.of(axis[0], axis[1])
    return Class.PeriodicBoundaryPolicy3D.of(axis[0], axis[1], axis[2])


# ------------------------------------------------------------------ writers
class Writer:
    """Writer(path; file_name, connect, digits, suffix) (Writer.jl:39-60):
    configuration files under <path>/config, raw files under <path>/raw."""

    def __init__(self, path: str, file_name: str = "result", connect: str = "_", digits: int = 4,
                 suffix: str = ".npz"):
        if suffix == ".jld2":
            warnings.warn("JLD2 needs an HDF5 library, absent here: raw files are written as .npz archives "
                          "of the same datasets", stacklevel=2)
            suffix = ".npz"
        if suffix != ".npz":
            raise ValueError(f"raw files are .npz archives, not {suffix!r}")
        self.path_ = path
        self.config_path_ = os.path.join(path, "config")
        self.raw_path_ = os.path.join(path, "raw")
        self.file_name_, self.connect_, self.digits_, self.suffix_ = file_name, connect, int(digits), suffix

    def get_file_name(self, step: int) -> str:
        """RawWriter get_file_name (RawWriter.jl:23-28): <raw>/<file_name><connect><step padded><suffix>."""
        return os.path.join(self.raw_path_, f"{self.file_name_}{self.connect_}{int(step):0{self.digits_}d}{self.suffix_}")

    def get_path(self) -> str:
        return self.path_


def WriterFrom(config: dict) -> Writer:
    """DataIO.Writer(config_dict) (Input.jl:106-114)."""
    w = config["writer"]
    return Writer(w["path"], file_name=w["file_name"], connect=w["connect"], digits=int(w["digits"]),
                  suffix=w["suffix"])


def appendix() -> OrderedDict:
    """RawWriter appendix() (RawWriter.jl:19-21)."""
    return OrderedDict([("TMSTEP", 0), ("TimeValue", 0.0), ("TimeStamp", ""), ("WriteStep", 0)])


def set_tables(host: Class.HostParticleSystem, int_properties, float_properties) -> None:
    """Class.set_int! / set_float! / set_n_particles! / set_is_alive! of a
    host system: rows 0..n-1 take the given tables, and are alive."""
    ip = np.asarray(int_properties, np.int32)
    fp = np.asarray(float_properties, np.float32)
    n = ip.shape[0]
    if fp.shape[0] != n:
        raise ValueError("int and float tables hold different numbers of rows")
    if ip.shape[1] != host.int_properties_.shape[1] or fp.shape[1] != host.float_properties_.shape[1]:
        raise ValueError("the tables' columns do not match the system's named index")
    if n > host.n_capacity:
        raise ValueError(f"{n} rows exceed the capacity {host.n_capacity}")
    host.int_properties_[:n] = ip
    host.float_properties_[:n] = fp
    host.set_n_particles(n)
    host.is_alive_[:] = 0
    host.set_is_alive()


def save(writer: Writer, host: Class.HostParticleSystem, app: dict, fmt: str = "%Y_%m_%d_%H_%M_%S") -> str:
    """save! (Writer.jl:156-175, RawWriter.jl:35-56): the alive rows of the
    host tables and the appendix (with a fresh TimeStamp) into the raw file
    of appendix["WriteStep"].  Returns the file name."""
    os.makedirs(writer.raw_path_, exist_ok=True)
    app = OrderedDict(app)
    app["TimeStamp"] = datetime.datetime.now().strftime(fmt)
    n = host.n_particles_
    mask = host.is_alive_[:n] == 1
    arrays = {"raw/int_properties": np.ascontiguousarray(host.int_properties_[:n][mask]),
              "raw/float_properties": np.ascontiguousarray(host.float_properties_[:n][mask])}
    for k, v in app.items():
        arrays[f"appendix/{k}"] = np.asarray(v)
    name = writer.get_file_name(app["WriteStep"])
    with open(name, "wb") as f:
        np.savez(f, **arrays)
    return name


def syncsave(writer: Writer, host: Class.HostParticleSystem, device: Class.ParticleSystem, app: dict) -> str:
    """syncsave! (Writer.jl:177-197): device state to the host mirror, then save."""
    Class.syncto(host, device)
    return save(writer, host, app)


def load_into(writer: Writer, host: Class.HostParticleSystem, step: int, app: dict | None = None) -> dict:
    """load! (Writer.jl:221-237): the raw file of `step` into the host
    system (rows alive) and its appendix values into `app`."""
    app = appendix() if app is None else app
    name = writer.get_file_name(step)
    ip, fp = _raw_tables(name)
    with np.load(name, allow_pickle=False) as z:
        for k in list(app.keys()):
            v = z[f"appendix/{k}"]
            app[k] = v.item() if v.ndim == 0 else v
    set_tables(host, ip, fp)
    return app


# ----------------------------------------------------------------- Output.jl
def config(parallel: Environment.Parallel, domain: Class.Domain2D, particle_system, neighbour_system,
           writer: Writer, config_dict: dict | None = None) -> OrderedDict:
    """DataIO.config (Output.jl:11-128): a configuration describing the
    handles, on top of config_dict (default: template())."""
    c = copy.deepcopy(config_dict if config_dict is not None else template())
    c["parallel"]["int"], c["parallel"]["float"] = "Int32", "Float32"
    c["parallel"]["backend"] = "rocm"
    c["parallel"]["device"] = int(parallel.device) + 1
    g = domain.args  # (gap, first_x, first_y, last_x, last_y) as given
    c["domain"].update(OrderedDict([("dimension", 2), ("gap", float(g[0])), ("first_x", float(g[1])),
                                    ("first_y", float(g[2])), ("last_x", float(g[3])), ("last_y", float(g[4]))]))
    ps = particle_system
    n, cap = int(ps.n_particles_), int(ps.n_capacity)
    assert n <= cap
    p = c["particle_system"]
    p["n_particles"] = n
    p["capacity_expand"] = f"n -> n + {cap - n}"
    p["int_named_tuple"] = OrderedDict(ps.named_index_.int_table.capacity)
    p["float_named_tuple"] = OrderedDict(ps.named_index_.float_table.capacity)
    p["basic_parameters"] = OrderedDict((k, v.item() if hasattr(v, "item") else v)
                                        for k, v in ps.basic_parameters_.items()
                                        if isinstance(v, (int, float, np.integer, np.floating)))
    ns = neighbour_system
    nc = c["neighbour_system"]
    nc["active_pair"] = [[int(a), int(b)] for a, b in ns.active_pair_.pair_vector_]
    nc["max_neighbour_number"] = int(ns.max_neighbour_number)
    policy = ns.periodic_boundary_.policy  # Output.jl:86-97
    if policy is Class.NonePeriodicBoundaryPolicy:
        nc["periodic_boundary_policy"]["type"] = "none"
    else:
        nc["periodic_boundary_policy"]["type"] = "2D" if len(policy.axis) == 2 else "3D"
        for k, a in enumerate(policy.axis):
            nc["periodic_boundary_policy"]["axis"][k] = a
    w = c["writer"]
    w["path"], w["file_name"], w["connect"] = writer.path_, writer.file_name_, writer.connect_
    w["digits"], w["suffix"] = writer.digits_, writer.suffix_
    return c


def writeConfig(writer: Writer, config_dict: dict, fmt: str = "json") -> str:
    """writeConfig (Output.jl:130-151): <path>/config/config.json (or .yaml)."""
    os.makedirs(writer.config_path_, exist_ok=True)
    if fmt == "yaml":
        import yaml
        name = os.path.join(writer.config_path_, "config.yaml")
        with open(name, "w") as f:
            yaml.safe_dump(json.loads(json.dumps(config_dict)), f, sort_keys=False)
        return name
    if fmt != "json":
        warnings.warn(f"unsupported format {fmt!r}, writing json", stacklevel=2)
    name = os.path.join(writer.config_path_, "config.json")
    with open(name, "w") as f:
        json.dump(config_dict, f, indent=4)
    return name
