"""Python binding of the MI355X HEOM library (libheom_hip.so, C ABI in include/heom_hip.h).

The binding is a thin ctypes layer: every call goes to the HIP library, which fails
loudly (HeomError) when it is missing or when no GPU is present -- there is no CPU
fallback.  ``Solver`` mirrors the reference's solver interface for this path
(population_dynamics_solver: reset / step_forward / hierarchy_top / hierarchy_norm /
hierarchy_mmult_left, dm-heom/include/heom/population_dynamics_solver.hpp and
dm-heom/include/heom/solver_base.hpp) on top of one library handle.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT, "lib", "libheom_hip.so")
HOST_LIB_PATH = os.path.join(ROOT, "lib", "libheom_host.so")
HEADER = os.path.join(os.path.dirname(ROOT), "include", "heom_hip.h")
HOST_HEADER = os.path.join(os.path.dirname(ROOT), "include", "heom_host.h")
CONFIG_DIR = os.path.join(ROOT, "configs")

HEOM_OK = 0
HEOM_E_INVALID = -1
HEOM_E_UNSUPPORTED = -2
HEOM_E_DEVICE = -3
HEOM_E_MEMORY = -4
HEOM_PRECISION_FP64 = 0
HEOM_PRECISION_FP32 = 1

# every function include/heom_hip.h declares
EXPORTED = [
    "heom_create", "heom_create_batch", "heom_instances", "heom_destroy", "heom_last_error", "heom_abi_version",
    "heom_set_hamiltonian", "heom_set_mask", "heom_get_mask", "heom_set_state", "heom_set_top",
    "heom_clear_state",
    "heom_get_state", "heom_get_top", "heom_hierarchy_norm", "heom_rhs", "heom_step_rk4",
    "heom_propagate", "heom_mmult_left", "heom_sync", "heom_timing", "heom_device_bytes",
    "heom_rk4_step_bytes", "heom_stage_kernel", "heom_propagate_observe", "heom_use_graphs",
    "heom_adjoint_forward", "heom_adjoint_backward", "heom_rhs_adjoint", "heom_autotune", "heom_compact",
    "heom_propagate_timedep", "heom_create_tss", "heom_tss_step", "heom_get_ados", "heom_observe",
    "heom_create_tss_batch", "heom_set_stage_config", "heom_propagate_observe_noalloc",
    "heom_add_pre_evaluate", "heom_clear_pre_evaluate", "heom_evaluate", "heom_lincomb", "heom_swap_buffers",
    "heom_add_buffers", "heom_buffer_count", "heom_set_buffer", "heom_get_buffer",
]
HEOM_NO_BUFFER = -1
HEOM_MAX_BUFFERS = 16
# int (*)(heom_handle*, double time, double step_size, int in, int out, void* user)
PRE_EVALUATE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_void_p)
HEOM_STAGE_TILE = 0
HEOM_STAGE_PIPELINED = 1
HEOM_STAGE_WIDE = 2
STAGE_KINDS = {HEOM_STAGE_TILE: "tile", HEOM_STAGE_PIPELINED: "pipelined", HEOM_STAGE_WIDE: "wide"}


class HeomError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"heom error {code}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_int32), ("matsubaras", ctypes.c_int32), ("baths", ctypes.c_int32),
        ("baths_max_per_state", ctypes.c_int32), ("width", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("matrices", ctypes.c_int64),
        ("coupling", ctypes.c_void_p), ("hamiltonian", ctypes.c_void_p),
        ("ado_tuples", ctypes.c_void_p), ("ado_plus", ctypes.c_void_p), ("ado_minus", ctypes.c_void_p),
        ("pf_same_sum", ctypes.c_void_p), ("pf_same", ctypes.c_void_p), ("pf_circ", ctypes.c_void_p),
        ("cbk", ctypes.c_void_p),
    ]


class Gradients(ctypes.Structure):
    """heom_gradients: output pointers (NULL skips an output)."""
    _fields_ = [("hamiltonian", ctypes.c_void_p), ("pf_same_sum", ctypes.c_void_p), ("pf_same", ctypes.c_void_p),
                ("pf_circ", ctypes.c_void_p), ("cbk", ctypes.c_void_p), ("state", ctypes.c_void_p)]


class TuneResult(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("xcd_chunk", ctypes.c_int32), ("workgroups", ctypes.c_int32),
                ("ados_per_workgroup", ctypes.c_int32), ("chosen", ctypes.c_int32), ("tile_split", ctypes.c_int32),
                ("ms_per_step", ctypes.c_double)]


class TimingInfo(ctypes.Structure):
    _fields_ = [("rk4_steps", ctypes.c_int64), ("rk4_seconds", ctypes.c_double),
                ("rhs_launches", ctypes.c_int64), ("last_call_seconds", ctypes.c_double)]


_lib = None


def build() -> str:
    import subprocess

    subprocess.run(["make", "-s", "-C", ROOT], check=True)
    return LIB_PATH


def load():
    """Load libheom_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HeomError(HEOM_E_UNSUPPORTED, f"libheom_hip.so not built: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        sig = {
            "heom_create": [ctypes.POINTER(Desc), ctypes.POINTER(vp)],
            "heom_create_batch": [ctypes.POINTER(Desc), ctypes.c_int32, ctypes.POINTER(vp)],
            "heom_instances": [vp, ctypes.POINTER(ctypes.c_int32)],
            "heom_destroy": [vp],
            "heom_abi_version": [],
            "heom_set_hamiltonian": [vp, vp],
            "heom_set_mask": [vp, vp],
            "heom_get_mask": [vp, vp],
            "heom_set_state": [vp, vp],
            "heom_set_top": [vp, vp],
            "heom_clear_state": [vp],
            "heom_get_state": [vp, vp],
            "heom_get_top": [vp, vp],
            "heom_hierarchy_norm": [vp, dbl, vp],
            "heom_rhs": [vp, vp],
            "heom_step_rk4": [vp, dbl, i64],
            "heom_propagate": [vp, dbl, i64, i64, vp],
            "heom_mmult_left": [vp, vp],
            "heom_sync": [vp],
            "heom_timing": [vp, ctypes.POINTER(TimingInfo)],
            "heom_device_bytes": [vp, ctypes.POINTER(i64)],
            "heom_rk4_step_bytes": [vp, ctypes.POINTER(i64)],
            "heom_stage_kernel": [vp, ctypes.POINTER(ctypes.c_int)],
            "heom_propagate_observe": [vp, dbl, i64, i64, vp, vp, vp],
            "heom_propagate_observe_noalloc": [vp, dbl, i64, i64, vp, vp, vp],
            "heom_use_graphs": [vp, ctypes.c_int],
            "heom_adjoint_forward": [vp, dbl, i64, i64, i64, vp, vp],
            "heom_adjoint_backward": [vp, vp, ctypes.POINTER(Gradients)],
            "heom_rhs_adjoint": [vp, vp],
            "heom_autotune": [vp, ctypes.c_int32, ctypes.POINTER(TuneResult), ctypes.c_int32,
                              ctypes.POINTER(ctypes.c_int32)],
            "heom_set_stage_config": [vp, ctypes.POINTER(TuneResult)],
            "heom_compact": [vp, ctypes.POINTER(i64)],
            "heom_propagate_timedep": [vp, dbl, i64, i64, vp, vp],
            "heom_create_tss": [ctypes.POINTER(Desc), vp, vp, vp, ctypes.POINTER(ctypes.c_void_p)],
            "heom_tss_step": [vp, i64],
            "heom_create_tss_batch": [vp, ctypes.c_int32, vp, vp, vp, ctypes.POINTER(ctypes.c_void_p)],
            "heom_get_ados": [vp, ctypes.c_int32, vp, i64, vp],
            "heom_observe": [vp, vp, vp, vp],
            "heom_add_pre_evaluate": [vp, PRE_EVALUATE_FN, vp],
            "heom_clear_pre_evaluate": [vp],
            "heom_evaluate": [vp, dbl, dbl, ctypes.c_int, ctypes.c_int, ctypes.c_int, dbl, ctypes.c_int, dbl,
                              ctypes.c_int],
            "heom_lincomb": [vp, ctypes.c_int, ctypes.c_int, vp, vp],
            "heom_swap_buffers": [vp, ctypes.c_int, ctypes.c_int],
            "heom_add_buffers": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
            "heom_buffer_count": [vp, ctypes.POINTER(ctypes.c_int)],
            "heom_set_buffer": [vp, ctypes.c_int, vp],
            "heom_get_buffer": [vp, ctypes.c_int, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.heom_last_error.argtypes = []
        L.heom_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != HEOM_OK:
        raise HeomError(rc, load().heom_last_error().decode())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128_flat(m) -> np.ndarray:
    m = np.asarray(m, dtype=np.complex128)
    return _f64(np.stack([m.real, m.imag], axis=-1).reshape(-1))


def _make_desc(states, matsubaras, baths, coupling, hamiltonian, tuples, plus, minus, pf_same_sum, pf_same,
               pf_circ, cbk, precision=HEOM_PRECISION_FP64):
    """A heom_desc over the reference instance arrays (returns the descriptor and the
    arrays it points into, which must stay alive until heom_create returns)."""
    coupling = np.ascontiguousarray(coupling, dtype=np.int32)
    keep = dict(
        coupling=coupling,
        hamiltonian=_f64(hamiltonian).reshape(-1),
        tuples=np.ascontiguousarray(tuples, dtype=np.int32).reshape(-1),
        plus=np.ascontiguousarray(plus, dtype=np.int32).reshape(-1),
        minus=np.ascontiguousarray(minus, dtype=np.int32).reshape(-1),
        pf_same_sum=_f64(pf_same_sum), pf_same=_f64(pf_same), pf_circ=_f64(pf_circ), cbk=_f64(cbk),
    )
    width = int(baths) * int(matsubaras)
    d = Desc()
    d.states, d.matsubaras, d.baths = int(states), int(matsubaras), int(baths)
    d.baths_max_per_state = int(coupling.shape[1]) if coupling.ndim == 2 else int(coupling.size // states)
    d.width = width
    d.precision = int(precision)
    d.matrices = int(keep["tuples"].size // max(width, 1))
    d.coupling = keep["coupling"].ctypes.data
    d.hamiltonian = keep["hamiltonian"].ctypes.data
    d.ado_tuples = keep["tuples"].ctypes.data
    d.ado_plus = keep["plus"].ctypes.data
    d.ado_minus = keep["minus"].ctypes.data
    d.pf_same_sum = keep["pf_same_sum"].ctypes.data
    d.pf_same = keep["pf_same"].ctypes.data
    d.pf_circ = keep["pf_circ"].ctypes.data
    d.cbk = keep["cbk"].ctypes.data
    return d, keep


def _instance_arrays(inst) -> dict:
    return dict(states=inst.states, matsubaras=inst.matsubaras, baths=inst.baths,
                coupling=np.asarray(inst.coupling).reshape(inst.states, -1), hamiltonian=inst.hamiltonian,
                tuples=inst.tuples, plus=inst.plus, minus=inst.minus, pf_same_sum=inst.pf_same_sum,
                pf_same=inst.pf_same, pf_circ=inst.pf_circ, cbk=inst.cbk)


class Handle:
    """One hierarchy, or a batch of hierarchies stepped together, on the GPU (owns all
    its device memory).  With a batch (count > 1) per-instance arrays carry a leading
    instance axis and per-ADO arrays are the instances' arrays one after the other."""

    count = 1

    @classmethod
    def from_desc(cls, desc: Desc, keepalive=None) -> "Handle":
        """Create from a filled descriptor (e.g. HostInstance.desc())."""
        return cls._create([desc], keepalive)

    @classmethod
    def _create(cls, descs, keepalive=None) -> "Handle":
        self = cls.__new__(cls)
        arr = (Desc * len(descs))(*descs)
        h = ctypes.c_void_p()
        _check(load().heom_create_batch(arr, len(descs), ctypes.byref(h)))
        self._h = h
        self.count = len(descs)
        # a batch may mix state counts and modes: arrays use the largest of each
        self.instance_states = [int(d.states) for d in descs]
        self.states = max(self.instance_states)
        self.instance_matrices = [int(d.matrices) for d in descs]
        self.matrices = sum(self.instance_matrices)
        self.width = max(int(d.width) for d in descs)
        self.baths = descs[0].baths
        self.matsubaras = descs[0].matsubaras
        self.precision = descs[0].precision
        self._keep = keepalive
        return self

    def __init__(self, *, states, matsubaras, baths, coupling, hamiltonian, tuples, plus, minus,
                 pf_same_sum, pf_same, pf_circ, cbk, precision=HEOM_PRECISION_FP64):
        d, keep = _make_desc(states, matsubaras, baths, coupling, hamiltonian, tuples, plus, minus, pf_same_sum,
                             pf_same, pf_circ, cbk, precision)
        other = Handle._create([d], keep)
        self.__dict__.update(other.__dict__)
        other._h = None

    @classmethod
    def from_instance(cls, inst, precision=HEOM_PRECISION_FP64) -> "Handle":
        """Create from any object carrying the reference instance arrays (states,
        matsubaras, baths, coupling, hamiltonian, tuples, plus, minus, pf_*, cbk)."""
        return cls(precision=precision, **_instance_arrays(inst))

    @classmethod
    def from_instances(cls, insts, precision=HEOM_PRECISION_FP64) -> "Handle":
        """A batch handle over several instances (any states, baths and bath parameters;
        arrays of the batch use the largest state count)."""
        made = [_make_desc(precision=precision, **_instance_arrays(i)) for i in insts]
        return cls._create([d for d, _ in made], [k for _, k in made])

    @classmethod
    def thermal_state_search(cls, desc: Desc, c, eigenvalues, epsilon, keepalive=None) -> "Handle":
        """heom_create_tss: a search handle for one instance (desc FP64), with C (the
        transposed eigenvectors of H), the eigenvalues (J) and epsilon."""
        self = cls.__new__(cls)
        N = int(desc.states)
        cc = _f64(np.stack([np.real(c), np.imag(c)], -1).reshape(-1))
        ev = _f64(np.stack([np.real(eigenvalues), np.imag(eigenvalues)], -1).reshape(-1))
        if cc.size != 2 * N * N or ev.size != 2 * N:
            raise ValueError("c must be N x N and eigenvalues of length N")
        eps = np.array([complex(epsilon).real, complex(epsilon).imag])
        h = ctypes.c_void_p()
        _check(load().heom_create_tss(ctypes.byref(desc), cc.ctypes.data, ev.ctypes.data, eps.ctypes.data,
                                      ctypes.byref(h)))
        self._h = h
        self.count = 1
        self.states = N
        self.instance_matrices = [int(desc.matrices)]
        self.matrices = int(desc.matrices)
        self.width = desc.width
        self.baths = desc.baths
        self.matsubaras = desc.matsubaras
        self.precision = desc.precision
        self._keep = keepalive
        return self

    @classmethod
    def thermal_state_search_batch(cls, descs, cs, eigenvalues, epsilon, keepalive=None) -> "Handle":
        """heom_create_tss_batch: one search handle over several instances (same states and
        baths), with one C and eigenvalue set per instance and a common epsilon."""
        self = cls.__new__(cls)
        n = len(descs)
        N = int(descs[0].states)
        cc = _f64(np.stack([np.stack([np.real(c), np.imag(c)], -1).reshape(-1) for c in cs]))
        ev = _f64(np.stack([np.stack([np.real(e), np.imag(e)], -1).reshape(-1) for e in eigenvalues]))
        if cc.size != n * 2 * N * N or ev.size != n * 2 * N:
            raise ValueError("one N x N C and N eigenvalues per instance")
        eps = np.array([complex(epsilon).real, complex(epsilon).imag])
        arr = (Desc * n)(*descs)
        h = ctypes.c_void_p()
        _check(load().heom_create_tss_batch(arr, n, cc.ctypes.data, ev.ctypes.data, eps.ctypes.data,
                                            ctypes.byref(h)))
        self._h = h
        self.count = n
        self.states = N
        self.instance_matrices = [int(d.matrices) for d in descs]
        self.matrices = sum(self.instance_matrices)
        self.width = descs[0].width
        self.baths = descs[0].baths
        self.matsubaras = descs[0].matsubaras
        self.precision = descs[0].precision
        self._keep = keepalive
        return self

    @classmethod
    def thermal_state_search_from_instances(cls, insts, setups) -> "Handle":
        """A batched search handle over instances' arrays with one setup (c, eigenvalues,
        epsilon) each."""
        made = [_make_desc(precision=HEOM_PRECISION_FP64, **_instance_arrays(i)) for i in insts]
        return cls.thermal_state_search_batch([d for d, _ in made], [s["c"] for s in setups],
                                              [s["eigenvalues"] for s in setups], setups[0]["epsilon"],
                                              [k for _, k in made])

    @classmethod
    def thermal_state_search_from_instance(cls, inst, setup: dict) -> "Handle":
        """A search handle over an instance's arrays with setup's c, eigenvalues, epsilon."""
        desc, keep = _make_desc(precision=HEOM_PRECISION_FP64, **_instance_arrays(inst))
        return cls.thermal_state_search(desc, setup["c"], setup["eigenvalues"], setup["epsilon"], keep)

    def tss_step(self, n: int = 1) -> None:
        """n iterations of the thermal state search (heom_tss_step)."""
        _check(load().heom_tss_step(self._h, int(n)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(load().heom_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # sizes
    @property
    def state_size(self) -> int:
        return self.matrices * self.states * self.states * 2

    # inputs
    def set_state(self, state) -> None:
        s = _f64(state)
        if s.size != self.state_size:
            raise ValueError("state has the wrong size")
        _check(load().heom_set_state(self._h, s.ctypes.data))

    def _per_instance(self, m) -> np.ndarray:
        """[count][N][N] complex from one N x N matrix (used for every instance), a stack,
        or a list with one matrix per instance of that instance's own size (padded with
        zeros to the batch's N)."""
        if isinstance(m, (list, tuple)) and len(m) == self.count and any(
                np.shape(x) != (self.states, self.states) for x in m):
            out = np.zeros((self.count, self.states, self.states), dtype=np.complex128)
            for k, x in enumerate(m):
                x = np.asarray(x)
                n = self.instance_states[k]
                if x.shape != (n, n):
                    raise ValueError(f"instance {k} expects a {n} x {n} matrix")
                out[k, :n, :n] = x
            return _c128_flat(out)
        m = np.asarray(m)
        if m.ndim == 2:
            m = np.broadcast_to(m, (self.count,) + m.shape)
        if m.shape != (self.count, self.states, self.states):
            raise ValueError("expected an N x N matrix or one per instance")
        return _c128_flat(m)

    def set_top(self, top) -> None:
        t = self._per_instance(top)
        _check(load().heom_set_top(self._h, t.ctypes.data))

    def clear_state(self) -> None:
        _check(load().heom_clear_state(self._h))

    def set_hamiltonian(self, H) -> None:
        h = self._per_instance(H)
        _check(load().heom_set_hamiltonian(self._h, h.ctypes.data))

    def set_mask(self, mask) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.size != self.matrices:
            raise ValueError("mask has the wrong size")
        _check(load().heom_set_mask(self._h, m.ctypes.data))

    def get_mask(self) -> np.ndarray:
        m = np.zeros(self.matrices, dtype=np.uint8)
        _check(load().heom_get_mask(self._h, m.ctypes.data))
        return m

    # outputs
    def get_state(self) -> np.ndarray:
        out = np.empty(self.state_size)
        _check(load().heom_get_state(self._h, out.ctypes.data))
        return out

    def observe(self, ops=None):
        """(populations [count, N], traces [count] complex or None) of the current state."""
        pops = np.zeros(self.count * self.states)
        if ops is None:
            _check(load().heom_observe(self._h, None, pops.ctypes.data, None))
            return pops.reshape(self.count, self.states), None
        o = self._per_instance(ops)
        tr = np.zeros(self.count * 2)
        _check(load().heom_observe(self._h, o.ctypes.data, pops.ctypes.data, tr.ctypes.data))
        return pops.reshape(self.count, self.states), tr[0::2] + 1j * tr[1::2]

    def get_ados(self, tuples, instance: int = 0) -> np.ndarray:
        """The ADOs with the given tuples ([n, W] ints) of one instance: complex [n, N, N]."""
        t = np.ascontiguousarray(tuples, dtype=np.int32)
        if t.ndim == 2 and t.shape[1] < self.width:  # an instance with fewer modes than the batch
            t = np.ascontiguousarray(np.pad(t, ((0, 0), (0, self.width - t.shape[1]))))
        if t.ndim != 2 or t.shape[1] != self.width:
            raise ValueError("tuples must be [n, W]")
        out = np.zeros(t.shape[0] * self.states * self.states * 2)
        _check(load().heom_get_ados(self._h, int(instance), t.ctypes.data, t.shape[0], out.ctypes.data))
        o = out.reshape(t.shape[0], self.states, self.states, 2)
        return o[..., 0] + 1j * o[..., 1]

    def get_top(self) -> np.ndarray:
        """The top ADO, (N, N); for a batch (count, N, N)."""
        out = np.empty(self.count * self.states * self.states * 2)
        _check(load().heom_get_top(self._h, out.ctypes.data))
        t = out.reshape(self.count, self.states, self.states, 2)
        t = t[..., 0] + 1j * t[..., 1]
        return t[0] if self.count == 1 else t

    def norms(self, threshold: float = -1.0) -> np.ndarray:
        out = np.empty(self.matrices)
        _check(load().heom_hierarchy_norm(self._h, float(threshold), out.ctypes.data))
        return out

    # evaluation
    def rhs(self) -> np.ndarray:
        out = np.empty(self.state_size)
        _check(load().heom_rhs(self._h, out.ctypes.data))
        return out

    def step_rk4(self, h: float, nsteps: int) -> None:
        _check(load().heom_step_rk4(self._h, float(h), int(nsteps)))

    # ---- per-evaluation interface (heom_evaluate and the buffer algebra) ----------------

    def evaluate(self, time: float, step_size: float, src: int = 0, out: int = HEOM_NO_BUFFER,
                 out_base: int = HEOM_NO_BUFFER, out_coeff: float = 1.0, acc: int = HEOM_NO_BUFFER,
                 acc_coeff: float = 0.0, acc_init: bool = False) -> None:
        """f = L(buffer src) after the pre-evaluate actions; buffer out = (buffer out_base
        or 0) + out_coeff f and buffer acc = (buffer src if acc_init else buffer acc) +
        acc_coeff f (ode::solve, dm-heom/src/heom/ode.cpp:97-196)."""
        _check(load().heom_evaluate(self._h, float(time), float(step_size), int(src), int(out), int(out_base),
                                    float(out_coeff), int(acc), float(acc_coeff), 1 if acc_init else 0))

    def lincomb(self, dst: int, terms) -> None:
        """buffer dst = sum of coeff * buffer src over terms [(coeff, src), ...] (1 to 8)."""
        terms = list(terms)
        src = np.array([t[1] for t in terms], dtype=np.int32)
        coeff = np.array([t[0] for t in terms], dtype=np.float64)
        _check(load().heom_lincomb(self._h, int(dst), len(terms), src.ctypes.data, coeff.ctypes.data))

    def swap_buffers(self, a: int, b: int) -> None:
        _check(load().heom_swap_buffers(self._h, int(a), int(b)))

    def add_buffers(self, n: int) -> int:
        """Adds n zeroed state buffers; returns the number of buffers now."""
        c = ctypes.c_int()
        _check(load().heom_add_buffers(self._h, int(n), ctypes.byref(c)))
        return c.value

    def buffer_count(self) -> int:
        c = ctypes.c_int()
        _check(load().heom_buffer_count(self._h, ctypes.byref(c)))
        return c.value

    def set_buffer(self, b: int, state) -> None:
        s = _f64(state)
        if s.size != self.state_size:
            raise ValueError(f"state has {s.size} values, expected {self.state_size}")
        _check(load().heom_set_buffer(self._h, int(b), s.ctypes.data))

    def get_buffer(self, b: int) -> np.ndarray:
        out = np.empty(self.state_size)
        _check(load().heom_get_buffer(self._h, int(b), out.ctypes.data))
        return out

    def add_pre_evaluate(self, fn) -> None:
        """ode::add_pre_evaluate_action: fn(handle, time, step_size, src, out) runs before
        every heom_evaluate (return None or 0; anything else fails the evaluation)."""
        def trampoline(_h, time, step_size, src, out, _user):
            try:
                rc = fn(self, time, step_size, src, out)
                return 0 if rc is None else int(rc)
            except Exception:  # an exception cannot cross the C boundary
                return -1
        cb = PRE_EVALUATE_FN(trampoline)
        self.__dict__.setdefault("_callbacks", []).append(cb)  # kept alive as long as the handle
        _check(load().heom_add_pre_evaluate(self._h, cb, None))

    def clear_pre_evaluate(self) -> None:
        _check(load().heom_clear_pre_evaluate(self._h))
        self._callbacks = []

    def propagate(self, h: float, nsteps: int, observe_steps: int) -> np.ndarray:
        """Tops every observe_steps steps: (rows, N, N); for a batch (rows, count, N, N)."""
        rows = nsteps // observe_steps + 1
        out = np.empty(rows * self.count * self.states * self.states * 2)
        _check(load().heom_propagate(self._h, float(h), int(nsteps), int(observe_steps), out.ctypes.data))
        t = out.reshape(rows, self.count, self.states, self.states, 2)
        t = t[..., 0] + 1j * t[..., 1]
        return t[:, 0] if self.count == 1 else t

    def propagate_timedep(self, h: float, nsteps: int, observe_steps: int, stage_h) -> np.ndarray:
        """propagate() with the Hamiltonian of every RK4 stage given: stage_h complex
        [nsteps, 4, N, N] (or [nsteps, 4, count, N, N]) in J, at the stage times t_n,
        t_n + h/2, t_n + h/2, t_n + h (heom_amd.field.stage_hamiltonians builds them)."""
        N = self.states
        sh = np.asarray(stage_h, dtype=np.complex128)
        if sh.ndim == 4:
            sh = np.broadcast_to(sh[:, :, None], (nsteps, 4, self.count, N, N))
        if sh.shape != (nsteps, 4, self.count, N, N):
            raise ValueError("stage_h must be [nsteps, 4, (count,) N, N]")
        flat = _c128_flat(sh)
        rows = nsteps // observe_steps + 1
        out = np.empty(rows * self.count * N * N * 2)
        _check(load().heom_propagate_timedep(self._h, float(h), int(nsteps), int(observe_steps), flat.ctypes.data,
                                             out.ctypes.data))
        t = out.reshape(rows, self.count, N, N, 2)
        t = t[..., 0] + 1j * t[..., 1]
        return t[:, 0] if self.count == 1 else t

    def propagate_observe(self, h: float, nsteps: int, observe_steps: int, ops=None, noalloc: bool = False):
        """Populations (rows, count, N) and, with ops (an N x N operator, or one per
        instance), traces Tr(ops sigma_0) (rows, count), reduced on the device.
        noalloc=True calls heom_propagate_observe_noalloc (no device allocation; the
        records must fit its fixed scratch)."""
        rows = nsteps // observe_steps + 1
        pops = np.empty(rows * self.count * self.states)
        traces = np.empty(rows * self.count * 2)
        o = self._per_instance(ops) if ops is not None else None
        fn = load().heom_propagate_observe_noalloc if noalloc else load().heom_propagate_observe
        _check(fn(self._h, float(h), int(nsteps), int(observe_steps), o.ctypes.data if o is not None else None,
                  pops.ctypes.data, traces.ctypes.data if o is not None else None))
        tr = traces.reshape(rows, self.count, 2)
        return pops.reshape(rows, self.count, self.states), (tr[..., 0] + 1j * tr[..., 1] if o is not None else None)

    def use_graphs(self, enable: bool = True) -> None:
        _check(load().heom_use_graphs(self._h, int(bool(enable))))

    def mmult_left(self, mx) -> None:
        m = self._per_instance(mx)
        _check(load().heom_mmult_left(self._h, m.ctypes.data))

    def sync(self) -> None:
        _check(load().heom_sync(self._h))

    def timing(self) -> TimingInfo:
        t = TimingInfo()
        _check(load().heom_timing(self._h, ctypes.byref(t)))
        return t

    def device_bytes(self) -> int:
        b = ctypes.c_int64()
        _check(load().heom_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def rk4_step_bytes(self) -> int:
        b = ctypes.c_int64()
        _check(load().heom_rk4_step_bytes(self._h, ctypes.byref(b)))
        return b.value

    def stage_kernel(self) -> str:
        """"pipelined", "tile" or "wide": the stage kernel this handle runs."""
        k = ctypes.c_int()
        _check(load().heom_stage_kernel(self._h, ctypes.byref(k)))
        return STAGE_KINDS[k.value]

    def compact(self) -> int:
        """Drop the masked ADOs from the device layout; returns the number kept."""
        n = ctypes.c_int64()
        _check(load().heom_compact(self._h, ctypes.byref(n)))
        return n.value

    def autotune(self, steps: int = 3) -> list:
        """Time every stage-kernel configuration on this handle's state (restored
        afterwards) and keep the fastest; one dict per candidate."""
        cap = 64
        res = (TuneResult * cap)()
        n = ctypes.c_int32()
        _check(load().heom_autotune(self._h, int(steps), res, cap, ctypes.byref(n)))
        return [{"kernel": STAGE_KINDS[r.kernel], "xcd_chunk": r.xcd_chunk, "workgroups": r.workgroups,
                 "ados_per_workgroup": r.ados_per_workgroup, "tile_split": r.tile_split, "chosen": bool(r.chosen),
                 "ms_per_step": r.ms_per_step}
                for r in res[:min(n.value, cap)]]

    def set_stage_config(self, config: dict) -> None:
        """Apply one configuration as returned by autotune() (kernel, xcd_chunk,
        workgroups, ados_per_workgroup) without timing anything."""
        kinds = {v: k for k, v in STAGE_KINDS.items()}
        r = TuneResult(kernel=kinds[config["kernel"]], xcd_chunk=int(config["xcd_chunk"]),
                       workgroups=int(config["workgroups"]), ados_per_workgroup=int(config["ados_per_workgroup"]),
                       tile_split=int(config.get("tile_split", 1)))
        _check(load().heom_set_stage_config(self._h, ctypes.byref(r)))

    # adjoint
    def rhs_adjoint(self) -> np.ndarray:
        """L^dag applied to the current state (reference layout)."""
        out = np.empty(self.state_size)
        _check(load().heom_rhs_adjoint(self._h, out.ctypes.data))
        return out

    def adjoint_forward(self, h: float, nsteps: int, observe_steps: int, ops, checkpoint_steps: int = 50):
        """Propagate from the current state recording X_rk = Tr(O_k sigma_0,k(t_r)) (rows,
        count) and checkpoints for adjoint_backward; the state ends at t = nsteps h."""
        rows = nsteps // observe_steps + 1
        o = self._per_instance(ops)
        out = np.empty(rows * self.count * 2)
        _check(load().heom_adjoint_forward(self._h, float(h), int(nsteps), int(observe_steps),
                                           int(checkpoint_steps), o.ctypes.data, out.ctypes.data))
        out = out.reshape(rows, self.count, 2)
        return out[..., 0] + 1j * out[..., 1]

    def adjoint_backward(self, weights, state: bool = True) -> dict:
        """Gradients of J = sum Re(w_rk X_rk) over the recorded sweep: complex arrays
        (dJ/dRe + i dJ/dIm) for the coefficient arrays -- hamiltonian (count, N, N),
        pf_same_sum (M), pf_same (B), pf_circ (B), cbk (B*K) -- and 'state', the real
        gradient with respect to the state the sweep started from."""
        w = np.asarray(weights, dtype=np.complex128)
        if w.ndim == 1:
            w = w[:, None]
        w = _c128_flat(w)
        N, B, K = self.states, self.baths, self.matsubaras
        bufs = dict(hamiltonian=np.zeros(self.count * N * N * 2), pf_same_sum=np.zeros(2 * self.matrices),
                    pf_same=np.zeros(2 * B), pf_circ=np.zeros(2 * B), cbk=np.zeros(2 * B * K))
        st = np.zeros(self.state_size) if state else None
        g = Gradients(*(bufs[k].ctypes.data for k in ("hamiltonian", "pf_same_sum", "pf_same", "pf_circ", "cbk")),
                      st.ctypes.data if st is not None else None)
        _check(load().heom_adjoint_backward(self._h, w.ctypes.data, ctypes.byref(g)))
        out = {k: v[0::2] + 1j * v[1::2] for k, v in bufs.items()}
        out["hamiltonian"] = out["hamiltonian"].reshape(self.count, N, N)
        out["state"] = st
        return out


# --------------------------------------------------------------------------------------
# host side: the config-file instance API (libheom_host.so, include/heom_host.h)
# --------------------------------------------------------------------------------------

TASK_POPULATION_DYNAMICS = 0
TASK_LINEAR_ABSORPTION = 1
TASK_THERMAL_STATE_SEARCH = 2

HOST_EXPORTED = [
    "heom_host_open", "heom_host_close", "heom_host_last_error", "heom_host_get_info", "heom_host_desc",
    "heom_host_top", "heom_host_cbk_sqrt", "heom_host_mask", "heom_host_dipoles", "heom_host_partitions",
    "heom_host_set_parameters", "heom_host_get_parameters", "heom_host_parameter_gradient",
    "heom_host_write_instance", "heom_host_read_instance_state", "heom_host_tss_setup",
    "heom_host_symmetric_eigen",
]


class HostInfo(ctypes.Structure):
    _fields_ = [("sites", ctypes.c_int64), ("states", ctypes.c_int64), ("baths", ctypes.c_int64),
                ("matsubaras", ctypes.c_int64), ("width", ctypes.c_int64), ("depth", ctypes.c_int64),
                ("matrices", ctypes.c_int64), ("first_non_plus", ctypes.c_int64),
                ("step_size", ctypes.c_double), ("steps", ctypes.c_int64),
                ("observe_steps", ctypes.c_int64), ("tensor_components", ctypes.c_int64),
                ("tss_delta", ctypes.c_double), ("tss_max_steps", ctypes.c_int64)]


class ConfigError(ValueError):
    """A config file the reference would reject (heom::config_error)."""


_host = None


def load_host():
    """Load libheom_host.so (raises if it has not been built)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise HeomError(HEOM_E_UNSUPPORTED, f"libheom_host.so not built: {HOST_LIB_PATH}")
        L = ctypes.CDLL(HOST_LIB_PATH)
        vp = ctypes.c_void_p
        sig = {
            "heom_host_open": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(vp)], ctypes.c_int),
            "heom_host_close": ([vp], ctypes.c_int),
            "heom_host_get_info": ([vp, ctypes.POINTER(HostInfo)], ctypes.c_int),
            "heom_host_desc": ([vp, ctypes.POINTER(Desc)], ctypes.c_int),
            "heom_host_top": ([vp, vp], ctypes.c_int),
            "heom_host_cbk_sqrt": ([vp, vp], ctypes.c_int),
            "heom_host_mask": ([vp, vp], ctypes.c_int),
            "heom_host_dipoles": ([vp, ctypes.c_int64, vp, vp], ctypes.c_int),
            "heom_host_partitions": ([ctypes.c_int, ctypes.c_int, vp, ctypes.c_int64], ctypes.c_int64),
            "heom_host_set_parameters": ([vp, vp, vp, vp, vp], ctypes.c_int),
            "heom_host_get_parameters": ([vp, vp, vp, vp, vp], ctypes.c_int),
            "heom_host_parameter_gradient": ([vp, ctypes.POINTER(Gradients), vp, vp, vp, vp], ctypes.c_int),
            "heom_host_write_instance": ([vp, ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
            "heom_host_tss_setup": ([vp, vp, vp, vp, vp], ctypes.c_int),
            "heom_host_symmetric_eigen": ([ctypes.c_int, vp, vp, vp], ctypes.c_int),
            "heom_host_read_instance_state": ([vp, ctypes.c_char_p, vp, vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        L.heom_host_last_error.argtypes = []
        L.heom_host_last_error.restype = ctypes.c_char_p
        _host = L
    return _host


def _host_check(rc: int) -> None:
    if rc != 0:
        msg = load_host().heom_host_last_error().decode()
        if rc == -5:
            raise ConfigError(msg)
        raise HeomError(rc, msg)


def symmetric_eigen(a) -> tuple:
    """(values, vectors as columns) of a real symmetric matrix: the host library's Jacobi
    routine of the thermal state search set-up."""
    m = _f64(a)
    n = int(np.sqrt(m.size))
    if n * n != m.size:
        raise ValueError("a must be square")
    w, v = np.zeros(n), np.zeros(n * n)
    _host_check(load_host().heom_host_symmetric_eigen(n, m.ctypes.data, w.ctypes.data, v.ctypes.data))
    return w, v.reshape(n, n)


def partitions(n: int, s: int) -> np.ndarray:
    """Integer partitions of n into at most s parts, in the reference's order."""
    L = load_host()
    cnt = L.heom_host_partitions(n, s, None, 0)
    if cnt < 0:
        _host_check(int(cnt))
    out = np.zeros((max(cnt, 1), s), dtype=np.int32)
    L.heom_host_partitions(n, s, out.ctypes.data, cnt)
    return out[:cnt]


class HostInstance:
    """A reference config file turned into kernel inputs by the C++ host code:
    config -> hierarchy_graph -> instance (population dynamics: identity mapping and
    rho_init on the top ADO; linear absorption: a ground state and |g><g|)."""

    def __init__(self, config_path: str, task: int = TASK_POPULATION_DYNAMICS, depth: int = 0):
        h = ctypes.c_void_p()
        _host_check(load_host().heom_host_open(str(config_path).encode(), int(task), int(depth), ctypes.byref(h)))
        self._p = h
        self.task = task
        info = HostInfo()
        _host_check(load_host().heom_host_get_info(h, ctypes.byref(info)))
        self.info = info
        self.states = int(info.states)
        self.sites = int(info.sites)
        self.baths = int(info.baths)
        self.matrices = int(info.matrices)
        self.width = int(info.width)

    def close(self) -> None:
        if getattr(self, "_p", None):
            load_host().heom_host_close(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def desc(self) -> Desc:
        d = Desc()
        _host_check(load_host().heom_host_desc(self._p, ctypes.byref(d)))
        return d

    def arrays(self) -> dict:
        """Copies of the kernel-input arrays (reference layout)."""
        d = self.desc()
        N, M, W, B, K = d.states, d.matrices, d.width, d.baths, d.matsubaras

        def arr(ptr, n, ctype, dtype):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,)).astype(dtype).copy()

        i32, f64 = ctypes.c_int32, ctypes.c_double
        out = dict(
            coupling=arr(d.coupling, N * d.baths_max_per_state, i32, np.int32),
            hamiltonian=arr(d.hamiltonian, 2 * N * N, f64, np.float64),
            tuples=arr(d.ado_tuples, M * W, i32, np.int32),
            plus=arr(d.ado_plus, M * W, i32, np.int32),
            minus=arr(d.ado_minus, M * W, i32, np.int32),
            pf_same_sum=arr(d.pf_same_sum, 2 * M, f64, np.float64),
            pf_same=arr(d.pf_same, 2 * B, f64, np.float64),
            pf_circ=arr(d.pf_circ, 2 * B, f64, np.float64),
            cbk=arr(d.cbk, 2 * B * K, f64, np.float64),
        )
        cs = np.zeros(B * K)
        _host_check(load_host().heom_host_cbk_sqrt(self._p, cs.ctypes.data))
        out["cbk_sqrt"] = cs
        return out

    def top(self) -> np.ndarray:
        t = np.zeros(2 * self.states * self.states)
        _host_check(load_host().heom_host_top(self._p, t.ctypes.data))
        t = t.reshape(self.states, self.states, 2)
        return t[..., 0] + 1j * t[..., 1]

    def mask(self) -> np.ndarray:
        m = np.zeros(self.matrices, dtype=np.uint8)
        _host_check(load_host().heom_host_mask(self._p, m.ctypes.data))
        return m

    def dipoles(self, t: int):
        n = self.states
        dm, dp = np.zeros(2 * n * n), np.zeros(2 * n * n)
        _host_check(load_host().heom_host_dipoles(self._p, int(t), dm.ctypes.data, dp.ctypes.data))
        dm, dp = dm.reshape(n, n, 2), dp.reshape(n, n, 2)
        return dm[..., 0] + 1j * dm[..., 1], dp[..., 0] + 1j * dp[..., 1]

    def set_parameters(self, hamiltonian=None, lam=None, gamma=None, omega=None) -> None:
        """Replace the site Hamiltonian (sites x sites) and the bath lambda_b, gamma_b =
        hbar / invnu_b and Omega_b, all in cm^-1 (None keeps); the prefactors are
        recomputed and the arrays desc() points to hold the new values."""
        H = _c128_flat(np.asarray(hamiltonian).reshape(self.sites, self.sites)) if hamiltonian is not None else None
        arrs = [_f64(np.asarray(v).reshape(self.baths)) if v is not None else None for v in (lam, gamma, omega)]
        _host_check(load_host().heom_host_set_parameters(
            self._p, H.ctypes.data if H is not None else None,
            *(a.ctypes.data if a is not None else None for a in arrs)))

    def get_parameters(self) -> dict:
        """hamiltonian (sites x sites complex), lambda, gamma, omega (B), all in cm^-1."""
        H = np.zeros(2 * self.sites * self.sites)
        lam, gam, om = np.zeros(self.baths), np.zeros(self.baths), np.zeros(self.baths)
        _host_check(load_host().heom_host_get_parameters(self._p, H.ctypes.data, lam.ctypes.data, gam.ctypes.data,
                                                         om.ctypes.data))
        H = H.reshape(self.sites, self.sites, 2)
        return {"hamiltonian": H[..., 0] + 1j * H[..., 1], "lambda": lam, "gamma": gam, "omega": om}

    def parameter_gradient(self, grads: dict, instance: int = 0) -> dict:
        """Handle.adjoint_backward's gradients (of batch instance `instance`) as gradients
        with respect to the parameters of set_parameters, per cm^-1: hamiltonian (sites x
        sites, dJ/dRe + i dJ/dIm per entry), lambda, gamma, omega (B)."""
        H = _c128_flat(np.asarray(grads["hamiltonian"])[instance] if np.ndim(grads["hamiltonian"]) == 3
                       else grads["hamiltonian"])
        arrs = {k: _c128_flat(grads[k]) for k in ("pf_same_sum", "pf_same", "pf_circ", "cbk")}
        g = Gradients(H.ctypes.data, arrs["pf_same_sum"].ctypes.data, arrs["pf_same"].ctypes.data,
                      arrs["pf_circ"].ctypes.data, arrs["cbk"].ctypes.data, None)
        dH = np.zeros(2 * self.sites * self.sites)
        dl, dg, do = np.zeros(self.baths), np.zeros(self.baths), np.zeros(self.baths)
        _host_check(load_host().heom_host_parameter_gradient(self._p, ctypes.byref(g), dH.ctypes.data,
                                                             dl.ctypes.data, dg.ctypes.data, do.ctypes.data))
        dH = dH.reshape(self.sites, self.sites, 2)
        return {"hamiltonian": dH[..., 0] + 1j * dH[..., 1], "lambda": dl, "gamma": dg, "omega": do}

    def tss_setup(self) -> dict:
        """heom_host_tss_setup: C, the eigenvalues (J), the Boltzmann top and epsilon."""
        N = self.states
        c, ev, top, eps = np.zeros(2 * N * N), np.zeros(2 * N), np.zeros(2 * N * N), np.zeros(2)
        _host_check(load_host().heom_host_tss_setup(self._p, c.ctypes.data, ev.ctypes.data, top.ctypes.data,
                                                    eps.ctypes.data))

        def cx(a, shape):
            return (a[0::2] + 1j * a[1::2]).reshape(shape)
        return {"c": cx(c, (N, N)), "eigenvalues": cx(ev, (N,)), "top": cx(top, (N, N)),
                "epsilon": complex(eps[0], eps[1])}

    def create_tss_handle(self) -> Handle:
        """heom_create_tss on the instance with its heom_host_tss_setup constants, the
        Boltzmann top loaded."""
        s = self.tss_setup()
        h = Handle.thermal_state_search(self.desc(), s["c"], s["eigenvalues"], s["epsilon"], keepalive=self)
        h.set_top(s["top"])
        return h

    def write_instance(self, path: str, state=None, top_only: bool = False, time: float = 0.0) -> None:
        """Write the reference's binary instance file with `state` (reference layout), or
        the top ADO only."""
        s = _f64(state) if state is not None else None
        if s is not None and s.size != self.matrices * self.states * self.states * 2:
            raise ValueError("state has the wrong size")
        _host_check(load_host().heom_host_write_instance(self._p, str(path).encode(),
                                                         s.ctypes.data if s is not None else None,
                                                         int(bool(top_only) or s is None), float(time)))

    def read_instance_state(self, path: str):
        """(state, time) from a binary instance file of this instance."""
        st = np.zeros(self.matrices * self.states * self.states * 2)
        t = ctypes.c_double()
        _host_check(load_host().heom_host_read_instance_state(self._p, str(path).encode(), st.ctypes.data,
                                                              ctypes.byref(t)))
        return st, t.value

    def create_handle(self, precision: int = HEOM_PRECISION_FP64) -> Handle:
        """heom_create on the instance, with its initial top ADO loaded."""
        d = self.desc()
        d.precision = int(precision)
        h = Handle.from_desc(d, keepalive=self)
        h.set_top(self.top())
        return h
