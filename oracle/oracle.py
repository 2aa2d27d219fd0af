"""CPU oracle harness for the HEOM population-dynamics / linear-absorption path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg import this module, as the checker; the product never does.

It binds ``heom_oracle.c`` (the FP64 statement-by-statement restatement of the
reference kernels, graph and precompute) through ctypes, and restates the host-side
pieces the reference keeps in C++:

* the INI config grammar of Boost.program_options plus the braced literals of the
  un-vendored noma/typa parser (``src/heom/config_base.cpp:39-63``); the exact typa
  grammar is parity unpinned, the subset used by ``data/examples_fmo/*.cfg`` is covered;
* ``config::check`` and the SI conversion (``src/heom/config.cpp:73-190``);
* the sites-to-states mappings (``src/heom/sites_to_states.cpp:16-213``);
* the instance assembly (``src/heom/instance.cpp:25-69, 483-495``);
* the dipole matrices of the linear-absorption pseudo pathway
  (``include/heom/dipole_matrix.hpp:30-74``, ``src/heom/dipole_matrix.cpp:182-208``,
  ``src/app_linear_absorption.cpp:88-163``), pinned with the absorption driver in
  tests/test_oracle_physics.py by the bath-free limit (the dipole correlation from
  scipy's expm and the config's directions), the exact second-cumulant line shape of a
  two-level monomer, and the rotating frame differing from the plain one by RK4
  truncation error only (falling as dt^4); planted sign errors in D+ or in the frame
  phase fail;
* the thermal state search set-up and loop
  (``include/heom/thermal_state_search_solver.hpp:124-325``,
  ``src/app_thermal_state_search.cpp:84-128``), with numpy's eigh in place of the
  un-vendored Eigen ComplexEigenSolver (pinned against the dense generator's null vector
  in tests/test_oracle_tss.py);
* the binary instance file (``src/heom/instance.cpp:281-377``), read and written with
  numpy independently of the product's C++ (no golden file exists in the reference; the
  layout is pinned by its field list and the size formula).
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libheom_oracle.so")

# include/heom/constants.hpp:21-24 and cl/heom/common.cl:45
HBAR = 1.05457148e-34
KB = 1.380650239684840458953781387048428305608616777679979316e-23
INVCM_TO_JOULE = 1.986456217327470903861498336971100097137908380913187891e-23

MASK_FALSE, MASK_TRUE, MASK_PROTECTED = 0, 1, 2

_lib_handle = None


class OracleError(RuntimeError):
    pass


class ConfigError(ValueError):
    """Mirrors heom::config_error (include/heom/config_error.hpp:17-23)."""


def build(force: bool = False) -> str:
    """Compile the C oracle in place (make -C oracle)."""
    import subprocess

    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise OracleError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        L.oracle_partition_count.restype = ctypes.c_int64
        L.oracle_partition_count.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_partitions.restype = ctypes.c_int64
        L.oracle_partitions.argtypes = [ctypes.c_int, ctypes.c_int, i32p]
        L.oracle_graph_nodes.restype = ctypes.c_int64
        L.oracle_graph_nodes.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_graph_first_non_plus.restype = ctypes.c_int64
        L.oracle_graph_first_non_plus.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_graph_build.restype = ctypes.c_int
        L.oracle_graph_build.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i64p]
        L.oracle_precompute.restype = None
        L.oracle_precompute.argtypes = [ctypes.c_int, ctypes.c_int, f64p, f64p, f64p, ctypes.c_double,
                                        ctypes.c_int64, ctypes.c_int, i32p, f64p, f64p, f64p, f64p, f64p]
        L.oracle_rhs.restype = None
        L.oracle_rhs.argtypes = [ctypes.c_void_p, f64p, f64p]
        L.oracle_rk4_step.restype = None
        L.oracle_rk4_step.argtypes = [ctypes.c_void_p, ctypes.c_double, f64p, f64p]
        L.oracle_propagate.restype = ctypes.c_int64
        L.oracle_propagate.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                       f64p, f64p, f64p]
        L.oracle_hierarchy_norm.restype = None
        L.oracle_hierarchy_norm.argtypes = [ctypes.c_int, ctypes.c_int64, f64p, ctypes.c_double, f64p,
                                            ctypes.c_void_p]
        L.oracle_mmult_left.restype = None
        L.oracle_mmult_left.argtypes = [ctypes.c_int, ctypes.c_int64, f64p, f64p]
        L.oracle_trace.restype = None
        L.oracle_trace.argtypes = [ctypes.c_int, ctypes.c_void_p, f64p, f64p]
        L.oracle_rhs_adjoint.restype = None
        L.oracle_rhs_adjoint.argtypes = [ctypes.c_void_p, f64p, f64p]
        L.oracle_adjoint_contract.restype = None
        L.oracle_adjoint_contract.argtypes = [ctypes.c_void_p, f64p, f64p, ctypes.c_double, f64p, f64p, f64p,
                                              f64p, f64p]
        L.oracle_adjoint_gradient.restype = ctypes.c_int
        L.oracle_adjoint_gradient.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                              ctypes.c_int64, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                                              f64p, ctypes.c_void_p]
        L.oracle_propagate_timedep.restype = ctypes.c_int64
        L.oracle_propagate_timedep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                               f64p, f64p, f64p, f64p]
        L.oracle_rhs_tss.restype = None
        L.oracle_rhs_tss.argtypes = [ctypes.c_void_p, f64p, f64p]
        L.oracle_mmult_right.restype = None
        L.oracle_mmult_right.argtypes = [ctypes.c_int, ctypes.c_int64, f64p, f64p]
        L.oracle_tss_step.restype = None
        L.oracle_tss_step.argtypes = [ctypes.c_void_p, f64p, f64p, ctypes.c_double, ctypes.c_double, f64p, f64p]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_num_threads.argtypes = []
        _lib_handle = L
        del u8p
    return _lib_handle


# --------------------------------------------------------------------------------------
# config grammar
# --------------------------------------------------------------------------------------

_NUM_RE = r"[+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)"


def _split_top(s: str) -> list[str]:
    """Split s at commas that are not nested in braces/parentheses."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "{(":
            depth += 1
        elif ch in "})":
            depth -= 1
            if depth < 0:
                raise ConfigError(f"unbalanced literal: {s!r}")
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if depth != 0:
        raise ConfigError(f"unbalanced literal: {s!r}")
    tail = "".join(cur).strip()
    if tail or out:
        out.append(tail)
    return out


def parse_literal(s: str):
    """Parse a noma/typa style literal: numbers, {..} lists (nested), (a, b) pairs."""
    s = s.strip()
    if s.startswith("{"):
        if not s.endswith("}"):
            raise ConfigError(f"malformed list literal: {s!r}")
        inner = s[1:-1].strip()
        if not inner:
            return []
        return [parse_literal(x) for x in _split_top(inner)]
    if s.startswith("("):
        if not s.endswith(")"):
            raise ConfigError(f"malformed pair literal: {s!r}")
        parts = _split_top(s[1:-1])
        if len(parts) != 2:
            raise ConfigError(f"pair literal needs two entries: {s!r}")
        return tuple(parse_literal(p) for p in parts)
    if re.fullmatch(_NUM_RE, s):
        if re.fullmatch(r"[+-]?\d+", s):
            return int(s)
        return float(s)
    return s


def read_ini(path: str) -> dict[str, str]:
    """Boost.program_options config-file reading: [section] prefixes, key=value,
    '#' comments; a repeated key is an error."""
    if not os.path.exists(path):
        raise ConfigError(f"could not open config file: {path}")
    out: dict[str, str] = {}
    section = ""
    with open(path) as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("[") and line.endswith("]"):
                section = line[1:-1].strip()
                continue
            if "=" not in line:
                raise ConfigError(f"invalid config line: {raw.rstrip()!r}")
            k, v = line.split("=", 1)
            key = (section + "." if section else "") + k.strip()
            if key in out:
                raise ConfigError(f"multiple occurrences of option '{key}'")
            out[key] = v.strip()
    return out


def _as_complex_matrix(v, name):
    if not isinstance(v, list) or not all(isinstance(r, list) for r in v):
        raise ConfigError(f"{name} must be a matrix literal")
    rows = []
    for r in v:
        row = []
        for x in r:
            if isinstance(x, tuple):
                row.append(complex(float(x[0]), float(x[1])))
            elif isinstance(x, (int, float)):
                row.append(complex(float(x), 0.0))
            else:
                raise ConfigError(f"{name}: bad entry {x!r}")
        rows.append(row)
    if len({len(r) for r in rows}) > 1:
        raise ConfigError(f"{name}: ragged matrix")
    return np.array(rows, dtype=np.complex128)


def _as_real_vector(v, name):
    if not isinstance(v, list):
        raise ConfigError(f"{name} must be a vector literal")
    try:
        return np.array([float(x) for x in v], dtype=np.float64)
    except (TypeError, ValueError):
        raise ConfigError(f"{name}: bad entry")


def _as_bool(s: str, name: str) -> bool:
    t = s.strip().lower()
    if t in ("true", "1", "yes", "on"):
        return True
    if t in ("false", "0", "no", "off"):
        return False
    raise ConfigError(f"faulty value: {name}={s!r}")


_REQUIRED_COMMON = [
    "program.task", "program.observations", "program.observe_steps",
    "filtering.strategy", "filtering.first_layer",
    "solver.stepper_type", "solver.step_size", "solver.steps", "solver.track_flows",
    "solver.flow_filename",
    "system.sites", "system.ado_depth", "system.hamiltonian",
    "baths.number", "baths.max_per_site", "baths.coupling", "baths.lambda", "baths.invnu",
    "baths.Omega", "baths.matsubaras", "baths.temperature",
]
_REQUIRED_TASK = {
    "population_dynamics": ["population_dynamics.rho_init"],
    "linear_absorption": ["dipole.directions", "dipole.strengths", "dipole.centers",
                          "dipole.tensor_prefactors", "dipole.tensor_components"],
    "thermal_state_search": ["thermal_state_search.delta", "thermal_state_search.max_steps"],
}


@dataclass
class Config:
    """Parsed and checked config in SI units (config::post_process)."""
    task: str
    observations: list
    observe_steps: int
    filtering_strategy: str
    filtering_first_layer: int
    stepper_type: str
    step_size: float
    steps: int
    track_flows: bool
    sites: int
    ado_depth: int
    hamiltonian: np.ndarray          # complex [sites, sites], J
    baths_number: int
    baths_max_per_site: int
    baths_coupling: np.ndarray       # int [sites, max_per_site]
    baths_lambda: np.ndarray         # J
    baths_invnu: np.ndarray          # s
    baths_omega: np.ndarray          # rad/s
    baths_matsubaras: int
    baths_temperature: float
    rho_init: np.ndarray | None = None
    dipole: dict = field(default_factory=dict)
    tss_delta: float = 0.0
    tss_max_steps: int = 0


def parse_config(path: str, task: str | None = None) -> Config:
    kv = read_ini(path)
    t = kv.get("program.task")
    if t is None:
        raise ConfigError("the option 'program.task' is required but missing")
    if task is not None and t != task:
        raise ConfigError("wrong program_task for this application")
    if t not in _REQUIRED_TASK:
        raise ConfigError(f"unsupported program.task: {t}")
    allowed = set(_REQUIRED_COMMON) | set(_REQUIRED_TASK[t])
    for k in kv:
        if k not in allowed:
            raise ConfigError(f"unrecognised option '{k}'")
    for k in sorted(allowed):
        if k not in kv:
            raise ConfigError(f"the option '{k}' is required but missing")

    lit = {k: parse_literal(v) for k, v in kv.items() if k not in ("solver.flow_filename",)}
    try:
        cfg = Config(
            task=t,
            observations=lit["program.observations"],
            observe_steps=int(lit["program.observe_steps"]),
            filtering_strategy=str(lit["filtering.strategy"]),
            filtering_first_layer=int(lit["filtering.first_layer"]),
            stepper_type=str(lit["solver.stepper_type"]),
            step_size=float(lit["solver.step_size"]),
            steps=int(lit["solver.steps"]),
            track_flows=_as_bool(kv["solver.track_flows"], "solver.track_flows"),
            sites=int(lit["system.sites"]),
            ado_depth=int(lit["system.ado_depth"]),
            hamiltonian=_as_complex_matrix(lit["system.hamiltonian"], "system.hamiltonian"),
            baths_number=int(lit["baths.number"]),
            baths_max_per_site=int(lit["baths.max_per_site"]),
            baths_coupling=np.array(lit["baths.coupling"], dtype=np.int64),
            baths_lambda=_as_real_vector(lit["baths.lambda"], "baths.lambda"),
            baths_invnu=_as_real_vector(lit["baths.invnu"], "baths.invnu"),
            baths_omega=_as_real_vector(lit["baths.Omega"], "baths.Omega"),
            baths_matsubaras=int(lit["baths.matsubaras"]),
            baths_temperature=float(lit["baths.temperature"]),
        )
    except (TypeError, ValueError) as e:
        raise ConfigError(f"faulty value: {e}")
    if t == "population_dynamics":
        cfg.rho_init = _as_complex_matrix(lit["population_dynamics.rho_init"], "population_dynamics.rho_init")
    elif t == "thermal_state_search":
        # thermal_state_search_config.cpp:19-23
        cfg.tss_delta = float(lit["thermal_state_search.delta"])
        cfg.tss_max_steps = int(lit["thermal_state_search.max_steps"])
    else:
        cfg.dipole = {
            "directions": np.array(lit["dipole.directions"], dtype=np.float64),
            "strengths": _as_real_vector(lit["dipole.strengths"], "dipole.strengths"),
            "centers": np.array(lit["dipole.centers"], dtype=np.float64),
            "tensor_prefactors": _as_real_vector(lit["dipole.tensor_prefactors"], "dipole.tensor_prefactors"),
            "tensor_components": np.array(lit["dipole.tensor_components"], dtype=np.int64),
        }
    check_config(cfg)
    post_process(cfg)
    return cfg


def check_config(c: Config) -> None:
    """config::check (src/heom/config.cpp:73-175) and the task checks
    (population_dynamics_config.cpp:310-319, dipole_config_entries.cpp:362-376)."""
    if not c.observations:
        raise ConfigError("observations must at least contain one entry.")
    if not c.observe_steps > 0:
        raise ConfigError("program.observe_steps must be > 0")
    if c.filtering_strategy not in ("none", "single_excitation"):
        raise ConfigError(f"unknown filtering.strategy: {c.filtering_strategy}")
    if c.filtering_strategy == "none" and c.filtering_first_layer != -1:
        raise ConfigError("filtering.first_layer must be -1, because it has no effect for filtering_strategy = none")
    if c.filtering_strategy == "single_excitation" and not c.filtering_first_layer >= 0:
        raise ConfigError("filtering.first_layer must be greater or equal to zero for filtering.strategy = single excitation filter")
    if c.ado_depth < c.filtering_first_layer:
        raise ConfigError("filtering.first_layer cannot be larger than system_ado_depth")
    if not c.step_size > 0.0:
        raise ConfigError("solver.step_size must be > 0")
    if not c.steps > 0:
        raise ConfigError("solver.steps must be > 0")
    if c.steps % c.observe_steps != 0:
        raise ConfigError("there's a remainder when dividing solver.steps by program.observe_steps")
    if not c.sites > 0:
        raise ConfigError("system.sites must be > 0")
    if not c.ado_depth > 0:
        raise ConfigError("system.ado_depth must be > 0")
    H = c.hamiltonian
    if H.shape != (c.sites, c.sites):
        raise ConfigError("system.hamiltonien must be a system.sites by system.sites matrix")
    if not np.array_equal(H, H.conj().T):
        raise ConfigError("system.hamiltonien must be hermitian")
    if not c.baths_number > 0:
        raise ConfigError("baths.number must be > 0")
    if not c.baths_max_per_site > 0:
        raise ConfigError("baths.max_per_site must be > 0")
    if not c.baths_max_per_site <= c.baths_number:
        raise ConfigError("baths.max_per_site must be <= baths.number")
    if c.baths_coupling.ndim != 2 or c.baths_coupling.shape != (c.sites, c.baths_max_per_site):
        raise ConfigError("baths.coupling must be a system.sites by baths_max_per_site_ matrix")
    if np.any(c.baths_coupling >= c.baths_number):
        raise ConfigError("baths.coupling values must between 0 and baths.number, or negative for non-coupling")
    if len(c.baths_lambda) != c.baths_number:
        raise ConfigError("baths.lambda must have a length of baths.number")
    if len(c.baths_invnu) != c.baths_number:
        raise ConfigError("baths.invnu must have a length of baths.number")
    if np.any(~(c.baths_invnu > 0)):
        raise ConfigError("baths.invnu must have values > 0")
    if len(c.baths_omega) != c.baths_number:
        raise ConfigError("baths.Omega must have a length of baths.number")
    if not 1 <= c.baths_matsubaras <= 10:
        raise ConfigError("baths.matsubaras must be between 1 and 10 (Pade table size)")
    if c.task == "population_dynamics":
        if c.rho_init.shape != (c.sites, c.sites):
            raise ConfigError("population_dynamics.rho_init must be a system.sites by system.sites matrix")
    elif c.task == "thermal_state_search":
        if c.step_size != 1.0:  # thermal_state_search_config.cpp:37-38
            raise ConfigError("solver.step_size_ must equal 1.0")
    else:
        d = c.dipole
        if not (len(d["directions"]) == len(d["strengths"]) == len(d["centers"])):
            raise ConfigError("dipole.directions, dipole.strengths and dipole.centers must have the same size")
        if len(d["tensor_prefactors"]) != len(d["tensor_components"]):
            raise ConfigError("dipole.tensor_prefactors and dipole.tensor_components must have the same size")


def post_process(c: Config) -> None:
    """Scale to SI units (src/heom/config.cpp:182-190)."""
    c.hamiltonian = c.hamiltonian * INVCM_TO_JOULE
    c.baths_lambda = c.baths_lambda * INVCM_TO_JOULE
    c.baths_omega = c.baths_omega * (INVCM_TO_JOULE / HBAR)
    c.baths_invnu = c.baths_invnu * 1.0e-15


# --------------------------------------------------------------------------------------
# graph, sites-to-states, instance
# --------------------------------------------------------------------------------------

@dataclass
class Graph:
    width: int
    depth: int
    tuples: np.ndarray     # int32 [M, W]
    plus: np.ndarray       # int32 [M, W]
    minus: np.ndarray      # int32 [M, W]
    nodes_per_depth: np.ndarray
    first_non_plus: int

    @property
    def nodes(self) -> int:
        return int(self.tuples.shape[0])


def partitions(n: int, s: int) -> np.ndarray:
    L = lib()
    cnt = L.oracle_partition_count(n, s)
    out = np.zeros((max(cnt, 1), s), dtype=np.int32)
    got = L.oracle_partitions(n, s, out)
    if got != cnt:
        raise OracleError("partition count mismatch")
    return out[:cnt]


def build_graph(width: int, depth: int) -> Graph:
    """hierarchy_graph(W, D) (src/heom/hierarchy_graph.cpp:30-181)."""
    L = lib()
    M = L.oracle_graph_nodes(width, depth)
    tuples = np.zeros((M, width), dtype=np.int32)
    plus = np.zeros((M, width), dtype=np.int32)
    minus = np.zeros((M, width), dtype=np.int32)
    npd = np.zeros(depth + 1, dtype=np.int64)
    rc = L.oracle_graph_build(width, depth, tuples, plus, minus, npd)
    if rc != 0:
        raise OracleError(f"oracle_graph_build failed: {rc}")
    return Graph(width, depth, tuples, plus, minus, npd, int(L.oracle_graph_first_non_plus(width, depth)))


def sites_to_states(sites: int, mode: str) -> int:
    if mode == "identity":
        return sites
    if mode == "with_ground_state":
        return sites + 1
    raise OracleError(f"unsupported sites_to_states mode {mode}")


def state_hamiltonian(H: np.ndarray, mode: str) -> np.ndarray:
    """src/heom/sites_to_states.cpp:44-56, 115-129."""
    if mode == "identity":
        return H.copy()
    n = H.shape[0] + 1
    out = np.zeros((n, n), dtype=np.complex128)
    out[1:, 1:] = H
    return out


def state_baths_coupling(coupling: np.ndarray, mode: str) -> np.ndarray:
    """src/heom/sites_to_states.cpp:131-143, 198-213."""
    if mode == "identity":
        return coupling.astype(np.int32).copy()
    out = -np.ones((coupling.shape[0] + 1, coupling.shape[1]), dtype=np.int32)
    out[1:, :] = coupling
    return out


class _CInstance(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_int), ("matsubaras", ctypes.c_int),
        ("baths_max_per_state", ctypes.c_int), ("width", ctypes.c_int),
        ("matrices", ctypes.c_int64), ("first_non_plus", ctypes.c_int64),
        ("coupling", ctypes.c_void_p), ("hamiltonian", ctypes.c_void_p),
        ("tuples", ctypes.c_void_p), ("plus", ctypes.c_void_p), ("minus", ctypes.c_void_p),
        ("pf_same_sum", ctypes.c_void_p), ("pf_same", ctypes.c_void_p),
        ("pf_circ", ctypes.c_void_p), ("cbk", ctypes.c_void_p), ("mask", ctypes.c_void_p),
    ]


class Instance:
    """The kernel-layout arrays of heom::instance (include/heom/instance.hpp:119-171),
    in the reference's ADO order and AoS interleaved-complex layout."""

    def __init__(self, *, states, matsubaras, baths, coupling, hamiltonian, graph: Graph,
                 lambda_, invnu, omega, temperature):
        self.states = int(states)
        self.matsubaras = int(matsubaras)
        self.baths = int(baths)
        self.coupling = np.ascontiguousarray(coupling, dtype=np.int32)
        self.baths_max_per_state = int(self.coupling.shape[1])
        self.graph = graph
        self.width = graph.width
        self.depth = graph.depth
        self.matrices = graph.nodes
        self.first_non_plus = graph.first_non_plus
        self.hamiltonian = np.ascontiguousarray(
            np.stack([hamiltonian.real, hamiltonian.imag], axis=-1).reshape(-1), dtype=np.float64)
        self.tuples = np.ascontiguousarray(graph.tuples.reshape(-1))
        self.plus = np.ascontiguousarray(graph.plus.reshape(-1))
        self.minus = np.ascontiguousarray(graph.minus.reshape(-1))
        self.lambda_ = np.ascontiguousarray(lambda_, dtype=np.float64)
        self.invnu = np.ascontiguousarray(invnu, dtype=np.float64)
        self.omega = np.ascontiguousarray(omega, dtype=np.float64)
        self.temperature = float(temperature)
        M, B, K = self.matrices, self.baths, self.matsubaras
        self.pf_same_sum = np.zeros(2 * M)
        self.pf_same = np.zeros(2 * B)
        self.pf_circ = np.zeros(2 * B)
        self.cbk = np.zeros(2 * B * K)
        self.cbk_sqrt = np.zeros(B * K)
        lib().oracle_precompute(B, K, self.lambda_, self.invnu, self.omega, self.temperature,
                                M, self.width, self.tuples, self.pf_same_sum, self.pf_same,
                                self.pf_circ, self.cbk, self.cbk_sqrt)
        self.mask = np.full(M, MASK_TRUE, dtype=np.uint8)
        self._c = None

    # -- state helpers --------------------------------------------------------------
    @property
    def state_size(self) -> int:
        return self.matrices * self.states * self.states * 2

    def zero_state(self) -> np.ndarray:
        return np.zeros(self.state_size, dtype=np.float64)

    def state_with_top(self, top: np.ndarray) -> np.ndarray:
        """instance::null_hierarchy + set_hierarchy_top (instance.cpp:448-453, 483-495)."""
        y = self.zero_state()
        t = np.asarray(top, dtype=np.complex128).reshape(self.states, self.states)
        y[: 2 * self.states * self.states] = np.stack([t.real, t.imag], -1).reshape(-1)
        return y

    def cstruct(self):
        if self._c is None:
            c = _CInstance()
            c.states, c.matsubaras = self.states, self.matsubaras
            c.baths_max_per_state, c.width = self.baths_max_per_state, self.width
            c.matrices, c.first_non_plus = self.matrices, self.first_non_plus
            for name in ("coupling", "hamiltonian", "tuples", "plus", "minus", "pf_same_sum",
                         "pf_same", "pf_circ", "cbk", "mask"):
                setattr(c, name, getattr(self, name).ctypes.data)
            self._c = c
        return ctypes.byref(self._c)

    def hamiltonian_matrix(self) -> np.ndarray:
        h = self.hamiltonian.reshape(self.states, self.states, 2)
        return h[..., 0] + 1j * h[..., 1]

    def set_hamiltonian(self, H: np.ndarray) -> None:
        self.hamiltonian[:] = np.stack([H.real, H.imag], -1).reshape(-1)

    # -- kernels ---------------------------------------------------------------------
    def rhs(self, state: np.ndarray) -> np.ndarray:
        out = np.empty_like(state)
        lib().oracle_rhs(self.cstruct(), np.ascontiguousarray(state), out)
        return out

    def rk4_step(self, y: np.ndarray, h: float, work: np.ndarray | None = None) -> None:
        if work is None:
            work = np.empty(5 * self.state_size)
        lib().oracle_rk4_step(self.cstruct(), h, y, work)

    def propagate(self, y: np.ndarray, h: float, nsteps: int, observe_steps: int) -> np.ndarray:
        """Advance y in place; returns the top ADOs (complex [rows, N, N])."""
        rows = nsteps // observe_steps + 1
        tops = np.zeros(rows * self.states * self.states * 2)
        work = np.empty(5 * self.state_size)
        got = lib().oracle_propagate(self.cstruct(), h, nsteps, observe_steps, y, work, tops)
        t = tops[: got * self.states * self.states * 2].reshape(got, self.states, self.states, 2)
        return t[..., 0] + 1j * t[..., 1]

    def propagate_timedep(self, y: np.ndarray, h: float, nsteps: int, observe_steps: int,
                          stage_h: np.ndarray) -> np.ndarray:
        """propagate() with a time-dependent Hamiltonian: stage_h complex [nsteps, 4, N, N],
        the Hamiltonian at the stage times t, t + h/2, t + h/2, t + h of each step
        (the e_field pre-evaluate action, include/heom/e_field.hpp:64-87)."""
        n = self.states
        rows = nsteps // observe_steps + 1
        sh = np.asarray(stage_h, dtype=np.complex128).reshape(nsteps, 4, n, n)
        sha = np.ascontiguousarray(np.stack([sh.real, sh.imag], -1).reshape(-1))
        tops = np.zeros(rows * n * n * 2)
        work = np.empty(5 * self.state_size)
        got = lib().oracle_propagate_timedep(self.cstruct(), h, nsteps, observe_steps, y, work, sha, tops)
        t = tops[: got * n * n * 2].reshape(got, n, n, 2)
        return t[..., 0] + 1j * t[..., 1]

    def norms(self, state: np.ndarray, threshold: float = -1.0) -> np.ndarray:
        out = np.zeros(self.matrices)
        lib().oracle_hierarchy_norm(self.states, self.matrices, np.ascontiguousarray(state),
                                    threshold, out, None)
        return out

    def mmult_left(self, mx: np.ndarray, state: np.ndarray) -> None:
        m = np.ascontiguousarray(np.stack([mx.real, mx.imag], -1).reshape(-1), dtype=np.float64)
        lib().oracle_mmult_left(self.states, self.matrices, m, state)

    def top(self, state: np.ndarray) -> np.ndarray:
        n = self.states
        t = state[: 2 * n * n].reshape(n, n, 2)
        return t[..., 0] + 1j * t[..., 1]

    # -- adjoint -----------------------------------------------------------------------
    def mmult_right(self, mx: np.ndarray, state: np.ndarray) -> None:
        m = np.ascontiguousarray(np.stack([mx.real, mx.imag], -1).reshape(-1), dtype=np.float64)
        lib().oracle_mmult_right(self.states, self.matrices, m, state)

    def rhs_tss(self, state: np.ndarray) -> np.ndarray:
        """f of the thermal state search (HEOM_TSS_EQUATION) with this instance's
        Hamiltonian as the pseudo-Hamiltonian."""
        out = np.zeros_like(state)
        lib().oracle_rhs_tss(self.cstruct(), np.ascontiguousarray(state), out)
        return out

    def tss_step(self, y: np.ndarray, setup: dict, n: int = 1) -> None:
        """n iterations of the thermal state search in place (the instance's Hamiltonian
        must be the pseudo-Hamiltonian, thermal_state_search_solver.hpp:199-209)."""
        c = np.ascontiguousarray(np.stack([setup["c"].real, setup["c"].imag], -1).reshape(-1))
        ev = np.ascontiguousarray(np.stack([setup["eigenvalues"].real, setup["eigenvalues"].imag], -1).reshape(-1))
        eps = complex(setup["epsilon"])
        work = np.zeros_like(y)
        for _ in range(n):
            lib().oracle_tss_step(self.cstruct(), c, ev, eps.real, eps.imag, y, work)

    def rhs_adjoint(self, mu: np.ndarray) -> np.ndarray:
        """L^dag mu, the transpose of rhs() (heom_oracle.c: oracle_rhs_adjoint)."""
        out = np.empty_like(mu)
        lib().oracle_rhs_adjoint(self.cstruct(), np.ascontiguousarray(mu), out)
        return out

    def _grad_arrays(self):
        N, M, B, W = self.states, self.matrices, self.baths, self.width
        return (np.zeros(2 * N * N), np.zeros(2 * M), np.zeros(2 * B), np.zeros(2 * B), np.zeros(2 * W))

    @staticmethod
    def _as_grads(arrays, N):
        def c(a):
            return a[0::2] + 1j * a[1::2]
        h, pss, same, circ, cbk = arrays
        return {"hamiltonian": c(h).reshape(N, N), "pf_same_sum": c(pss), "pf_same": c(same),
                "pf_circ": c(circ), "cbk": c(cbk)}

    def adjoint_contract(self, mu: np.ndarray, y: np.ndarray, w: float = 1.0) -> dict:
        """w * d<mu, L y>/d(kappa) per coefficient array, complex dRe + i dIm."""
        arrays = self._grad_arrays()
        lib().oracle_adjoint_contract(self.cstruct(), np.ascontiguousarray(mu), np.ascontiguousarray(y), w,
                                      *arrays)
        return self._as_grads(arrays, self.states)

    def adjoint_gradient(self, y0: np.ndarray, h: float, nsteps: int, observe_steps: int, op: np.ndarray,
                         weights: np.ndarray, checkpoint_steps: int = 50):
        """J = sum_r Re(w_r Tr(O sigma_0(t_r))): the observables X_r (complex [rows]) and the
        gradients of J (dict of complex arrays, plus 'state', the real gradient with
        respect to y0), by the discrete adjoint of the RK4 trajectory."""
        rows = nsteps // observe_steps + 1
        N = self.states
        opa = np.ascontiguousarray(np.stack([np.real(op), np.imag(op)], -1).reshape(-1), dtype=np.float64)
        w = np.asarray(weights, dtype=np.complex128).reshape(rows)
        wa = np.ascontiguousarray(np.stack([w.real, w.imag], -1).reshape(-1))
        observed = np.zeros(2 * rows)
        arrays = self._grad_arrays()
        g_state = np.zeros(self.state_size)
        rc = lib().oracle_adjoint_gradient(self.cstruct(), h, nsteps, observe_steps, checkpoint_steps, opa, wa,
                                           np.ascontiguousarray(y0, dtype=np.float64), observed, *arrays,
                                           g_state.ctypes.data)
        if rc != 0:
            raise OracleError("oracle_adjoint_gradient failed")
        g = self._as_grads(arrays, N)
        g["state"] = g_state
        return observed[0::2] + 1j * observed[1::2], g

    def precomputed(self, lambda_: np.ndarray, invnu: np.ndarray, omega: np.ndarray) -> dict:
        """The coefficient arrays oracle_precompute gives for other bath parameters (SI)."""
        M, B, K = self.matrices, self.baths, self.matsubaras
        pss, same, circ, cbk, cs = np.zeros(2 * M), np.zeros(2 * B), np.zeros(2 * B), np.zeros(2 * B * K), \
            np.zeros(B * K)
        lib().oracle_precompute(B, K, np.ascontiguousarray(lambda_, dtype=np.float64),
                                np.ascontiguousarray(invnu, dtype=np.float64),
                                np.ascontiguousarray(omega, dtype=np.float64), self.temperature, M, self.width,
                                self.tuples, pss, same, circ, cbk, cs)

        def c(a):
            return a[0::2] + 1j * a[1::2]
        return {"pf_same_sum": c(pss), "pf_same": c(same), "pf_circ": c(circ), "cbk": c(cbk)}

    def bath_parameter_gradient(self, grads: dict, rel: float = 1e-5) -> np.ndarray:
        """dJ/d(lambda_b, gamma_b, Omega_b), rows in that order, all per cm^-1, from the
        coefficient gradients by central differences of oracle_precompute: lambda_b and
        Omega_b as in the config, gamma_b = hbar / invnu_b as an energy in cm^-1."""
        B = self.baths
        out = np.zeros((3, B))
        keys = ("pf_same_sum", "pf_same", "pf_circ", "cbk")
        base = np.stack([self.lambda_ / INVCM_TO_JOULE, HBAR / (self.invnu * INVCM_TO_JOULE),
                         self.omega * HBAR / INVCM_TO_JOULE])

        def si(p):
            return p[0] * INVCM_TO_JOULE, HBAR / (p[1] * INVCM_TO_JOULE), p[2] * INVCM_TO_JOULE / HBAR

        for b in range(B):
            for k in range(3):
                step = rel * max(abs(base[k, b]), 1.0)
                vals = []
                for sgn in (1.0, -1.0):
                    p = base.copy()
                    p[k, b] += sgn * step
                    vals.append(self.precomputed(*si(p)))
                out[k, b] = sum(float(np.sum(np.real(np.conj(grads[key]) * (vals[0][key] - vals[1][key]))))
                                for key in keys) / (2.0 * step)
        return out


def field_stage_hamiltonians(H0: np.ndarray, d_minus: np.ndarray, d_plus: np.ndarray, field, t0: float, h: float,
                             nsteps: int) -> np.ndarray:
    """[nsteps, 4, N, N]: H(t) = H0 - E(t) D- - conj(E(t)) D+ (include/heom/e_field.hpp:78)
    at the RK4 stage times t_n, t_n + h/2, t_n + h/2, t_n + h, t_n = t0 + n h."""
    c = np.array([0.0, 0.5, 0.5, 1.0])
    times = t0 + h * (np.arange(nsteps)[:, None] + c[None, :])
    E = np.vectorize(field, otypes=[np.complex128])(times)
    return H0[None, None] - E[..., None, None] * d_minus[None, None] - np.conj(E)[..., None, None] * d_plus[None, None]


# --------------------------------------------------------------------------------------
# thermal state search (include/heom/thermal_state_search_solver.hpp:124-300,
# src/app_thermal_state_search.cpp:84-128)
# --------------------------------------------------------------------------------------

TSS_GROUND_STATE_ENERGY = 100000.0  # include/heom/constants.hpp:27
TSS_EPSILON = 1.0e14                # thermal_state_search_solver.hpp:201-206 ("Version C")


def thermal_state_setup(H: np.ndarray, temperature: float, mode: str = "identity") -> dict:
    """thermal_state_search_solver::initialize (:124-209) for a real symmetric state
    Hamiltonian H (J): eigenvectors V (numpy's eigh in place of the un-vendored Eigen
    ComplexEigenSolver), C = V^T, the eigenvalues, the Boltzmann top
    sigma_0 = C^T diag(exp(-beta E) / Z) C and the constant epsilon of the
    pseudo-Hamiltonian.  With a ground state its energy is raised to keep it empty."""
    H = np.array(H, dtype=np.complex128)
    if mode == "with_ground_state":
        H[0, 0] = TSS_GROUND_STATE_ENERGY
    if np.any(H.imag != 0):
        raise ValueError("the thermal state search needs a real symmetric Hamiltonian")
    E, V = np.linalg.eigh(H.real)
    beta = 1.0 / (KB * temperature)
    p = np.exp(-(E - E.min()) * beta)
    p /= p.sum()
    C = V.T.astype(np.complex128)
    top = C.T @ np.diag(p) @ C
    N = H.shape[0]
    return {"c": C, "eigenvalues": E.astype(np.complex128), "top": top, "epsilon": complex(TSS_EPSILON),
            "pseudo_hamiltonian": np.full((N, N), TSS_EPSILON, dtype=np.complex128)}


def thermal_state_search(cfg: Config, graph: Graph | None = None, max_iterations: int | None = None):
    """app_thermal_state_search main loop (src/app_thermal_state_search.cpp:84-128): at
    least solver.steps / observe_steps iterations of observe_steps search steps, then
    more until the 1-norm of the top changes by less than delta across an iteration or
    max_steps / observe_steps iterations are done.  Returns (tops [iterations + 1, N, N],
    final state, iterations)."""
    inst = instance_from_config(cfg, "identity", graph)
    setup = thermal_state_setup(inst.hamiltonian_matrix(), cfg.baths_temperature)
    inst.set_hamiltonian(setup["pseudo_hamiltonian"])
    y = inst.state_with_top(setup["top"])
    min_it = cfg.steps // cfg.observe_steps
    max_it = max(min_it, cfg.tss_max_steps // cfg.observe_steps)
    if max_iterations is not None:
        max_it = min(max_it, max_iterations)
    tops = [inst.top(y)]
    it = 0
    while it < max_it:
        inst.tss_step(y, setup, cfg.observe_steps)
        tops.append(inst.top(y))
        it += 1
        if it > min_it and abs(np.abs(tops[-2]).sum() - np.abs(tops[-1]).sum()) < cfg.tss_delta:
            break
    return np.array(tops), y, it


# --------------------------------------------------------------------------------------
# binary instance file (src/heom/instance.cpp:281-377, include/heom/instance.hpp:31-211)
# --------------------------------------------------------------------------------------

# packed header: version, hierarchy_top_only, hierarchy_layout, hamiltonian_layout,
# sites_to_states_mode, time, first_non_plus_id, then seven int64 sizes
_INSTANCE_HEADER = np.dtype([
    ("version", "<i4"), ("top_only", "u1"), ("hierarchy_layout", "<i4"), ("hamiltonian_layout", "<i4"),
    ("sts_mode", "<i4"), ("time", "<f8"), ("first_non_plus", "<i4"),
    ("states", "<i8"), ("matrices", "<i8"), ("matsubaras", "<i8"), ("baths_per_state", "<i8"),
    ("baths", "<i8"), ("width", "<i8"), ("depth", "<i8"),
])
STS_MODES = {"identity": 0, "with_ground_state": 1, "excited_state_absorption": 2}


def instance_file_size(states: int, matrices: int, matsubaras: int, baths: int, width: int, top_only: bool) -> int:
    """Bytes of an instance file: the 85-byte header, then the buffers in file order."""
    nn2 = states * states * 2
    return (_INSTANCE_HEADER.itemsize + 8 * (nn2 if top_only else nn2 * matrices) + 8 * nn2
            + 4 * 3 * matrices * width + 8 * (2 * matrices + 4 * baths + 3 * baths * matsubaras))


def write_instance_file(path: str, inst: Instance, state: np.ndarray, top_only: bool = False, time: float = 0.0,
                        mode: str = "identity") -> None:
    """instance::write_file (src/heom/instance.cpp:333-377) for an oracle instance."""
    h = np.zeros((), dtype=_INSTANCE_HEADER)
    h["top_only"] = int(bool(top_only))
    h["sts_mode"] = STS_MODES[mode]
    h["time"] = time
    h["first_non_plus"] = inst.first_non_plus
    for k, v in (("states", inst.states), ("matrices", inst.matrices), ("matsubaras", inst.matsubaras),
                 ("baths_per_state", inst.baths_max_per_state), ("baths", inst.baths), ("width", inst.width),
                 ("depth", inst.depth)):
        h[k] = v
    st = np.asarray(state, dtype="<f8").reshape(-1)
    nn2 = inst.states * inst.states * 2
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(st[:nn2].tobytes() if top_only else st.tobytes())
        for a, t in ((inst.hamiltonian, "<f8"), (inst.tuples, "<i4"), (inst.plus, "<i4"), (inst.minus, "<i4"),
                     (inst.pf_same_sum, "<f8"), (inst.pf_same, "<f8"), (inst.pf_circ, "<f8"), (inst.cbk, "<f8"),
                     (inst.cbk_sqrt, "<f8")):
            f.write(np.ascontiguousarray(a, dtype=t).tobytes())


def read_instance_file(path: str) -> dict:
    """instance::read_file (src/heom/instance.cpp:281-330): header fields and buffers
    (hierarchy as written: the top ADO only for a top-only file)."""
    with open(path, "rb") as f:
        data = f.read()
    h = np.frombuffer(data, dtype=_INSTANCE_HEADER, count=1)[0]
    out = {k: h[k].item() for k in _INSTANCE_HEADER.names}
    out["top_only"] = bool(out["top_only"])
    N, M, K, B, W = (out[k] for k in ("states", "matrices", "matsubaras", "baths", "width"))
    if len(data) != instance_file_size(N, M, K, B, W, out["top_only"]):
        raise ValueError("instance file has the wrong size for its header")
    off = _INSTANCE_HEADER.itemsize
    nn2 = N * N * 2

    def take(n, t):
        nonlocal off
        a = np.frombuffer(data, dtype=t, count=n, offset=off).copy()
        off += a.nbytes
        return a

    out["hierarchy"] = take(nn2 if out["top_only"] else nn2 * M, "<f8")
    out["hamiltonian"] = take(nn2, "<f8")
    for k in ("tuples", "plus", "minus"):
        out[k] = take(M * W, "<i4")
    out["pf_same_sum"] = take(2 * M, "<f8")
    out["pf_same"] = take(2 * B, "<f8")
    out["pf_circ"] = take(2 * B, "<f8")
    out["cbk"] = take(2 * B * K, "<f8")
    out["cbk_sqrt"] = take(B * K, "<f8")
    return out


def trace(d: np.ndarray | None, top: np.ndarray) -> complex:
    n = top.shape[0]
    t = np.ascontiguousarray(np.stack([top.real, top.imag], -1).reshape(-1), dtype=np.float64)
    out = np.zeros(2)
    if d is None:
        lib().oracle_trace(n, None, t, out)
    else:
        dd = np.ascontiguousarray(np.stack([d.real, d.imag], -1).reshape(-1), dtype=np.float64)
        lib().oracle_trace(n, dd.ctypes.data, t, out)
    return complex(out[0], out[1])


def instance_from_config(cfg: Config, mode: str = "identity", graph: Graph | None = None) -> Instance:
    """heom::instance(config, sts_mode, graph) (src/heom/instance.cpp:25-69)."""
    if graph is None:
        graph = build_graph(cfg.baths_number * cfg.baths_matsubaras, cfg.ado_depth)
    return Instance(
        states=sites_to_states(cfg.sites, mode),
        matsubaras=cfg.baths_matsubaras,
        baths=cfg.baths_number,
        coupling=state_baths_coupling(cfg.baths_coupling, mode),
        hamiltonian=state_hamiltonian(cfg.hamiltonian, mode),
        graph=graph,
        lambda_=cfg.baths_lambda, invnu=cfg.baths_invnu, omega=cfg.baths_omega,
        temperature=cfg.baths_temperature,
    )


# --------------------------------------------------------------------------------------
# linear absorption (src/app_linear_absorption.cpp:88-163)
# --------------------------------------------------------------------------------------

def dipole_matrices(cfg: Config, tensor_index: int):
    """D- (pre-scaled by the tensor prefactor) and D+ = (D-)^T for the pseudo pathway
    +{(-, L), (+, L)} with a ground state (include/heom/dipole_pseudo_pathway.hpp:20-36,
    include/heom/dipole_matrix.hpp:30-74, src/heom/dipole_matrix.cpp:182-208)."""
    d = cfg.dipole
    comps = d["tensor_components"][tensor_index]
    mats = []
    for spec_index in (0, 1):
        comp = int(comps[spec_index])
        vec = np.array([d["directions"][i][comp] * d["strengths"][i] for i in range(cfg.sites)])
        n = cfg.sites + 1
        mat = np.zeros((n, n), dtype=np.complex128)
        mat[0, 1:] = vec
        if spec_index == 1:  # the second dipole of the pathway has sign plus: transpose
            mat = mat.T.copy()
        mats.append(mat)
    dm, dp = mats
    dm = dm * complex(float(d["tensor_prefactors"][tensor_index]), 0.0)
    return dm, dp


def rotating_frame_phase(e0: float, times: np.ndarray) -> np.ndarray:
    """exp(-i E0 t / hbar): the phase the excited block loses when E0 is taken out of it."""
    return np.exp(-1j * e0 * times / HBAR)


def linear_absorption(cfg: Config, nsteps: int | None = None, observe_steps: int | None = None,
                      rotating_frame: float = 0.0):
    """Averaged dipole correlation Tr(D- sigma_0(t)) over the tensor components.

    rotating_frame E0 (in J) subtracts E0 from the excited block of H and restores the
    phase exp(-i E0 t / hbar) analytically; 0 reproduces the reference propagation."""
    nsteps = cfg.steps if nsteps is None else nsteps
    observe_steps = cfg.observe_steps if observe_steps is None else observe_steps
    graph = build_graph(cfg.baths_number * cfg.baths_matsubaras, cfg.ado_depth)
    times = np.arange(nsteps // observe_steps + 1) * observe_steps * cfg.step_size
    acc = np.zeros(len(times), dtype=np.complex128)
    ncomp = len(cfg.dipole["tensor_prefactors"])
    for t_idx in range(ncomp):
        inst = instance_from_config(cfg, "with_ground_state", graph)
        if rotating_frame != 0.0:
            H = state_hamiltonian(cfg.hamiltonian, "with_ground_state")
            H[1:, 1:] -= rotating_frame * np.eye(cfg.sites)
            inst.set_hamiltonian(H)
        top0 = np.zeros((inst.states, inst.states), dtype=np.complex128)
        top0[0, 0] = 1.0
        y = inst.state_with_top(top0)
        dm, dp = dipole_matrices(cfg, t_idx)
        inst.mmult_left(dp, y)
        tops = inst.propagate(y, cfg.step_size, nsteps, observe_steps)
        tr = np.array([trace(dm, tops[r]) for r in range(tops.shape[0])])
        if rotating_frame != 0.0:
            tr = tr * rotating_frame_phase(rotating_frame, times)
        acc += tr
    return times, acc / ncomp
