"""Helpers shared by the tests: instances of the workload configs and comparisons."""
import os

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIG_DIR = os.path.join(ROOT, "dm-heom_amd", "configs")
FMO_DIR = os.path.join(ROOT, "tests", "golden", "fmo")


def config_path(name):
    return os.path.join(CONFIG_DIR, f"{name}.cfg")


def load_config(name):
    return O.parse_config(config_path(name))


def instance(name, depth=None, mode="identity"):
    cfg = load_config(name)
    if depth is not None:
        cfg.ado_depth = depth
    return cfg, O.instance_from_config(cfg, mode)


def random_state(inst, seed, scale_top=1.0):
    """A dense random hierarchy state whose ADO magnitudes fall off with depth."""
    rng = np.random.default_rng(seed)
    depth = inst.graph.tuples.sum(axis=1)
    amp = (0.3 ** depth)[:, None, None]
    n = inst.states
    z = (rng.normal(size=(inst.matrices, n, n)) + 1j * rng.normal(size=(inst.matrices, n, n))) * amp
    z[0] *= scale_top
    return np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1))


def as_complex(state, inst):
    s = state.reshape(inst.matrices, inst.states, inst.states, 2)
    return s[..., 0] + 1j * s[..., 1]


def rel_err(a, b, floor_frac=1e-12):
    """max |a - b| / max(|b|, floor_frac * max|b|), elementwise over complex arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), floor_frac * np.max(np.abs(b)))
    scale = np.where(scale == 0, 1.0, scale)
    return float(np.max(np.abs(a - b) / scale))


def max_abs_rel(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def make_instance(N, B, K, D, coupling, seed=0, complex_h=False, omega=None, temperature=300.0):
    """A seeded instance of N sites (H about 12,000-12,600 cm^-1, couplings up to 100 cm^-1),
    B Drude-Lorentz baths of K modes each (lambda 20-60 cm^-1, 1/gamma 30-120 fs), depth D;
    coupling[b] lists the states bath b couples to."""
    rng = np.random.default_rng(seed)
    H = np.diag(rng.uniform(12000, 12600, N)).astype(np.complex128)
    for i in range(N):
        for j in range(i + 1, N):
            v = rng.uniform(-100, 100) + (1j * rng.uniform(-40, 40) if complex_h else 0.0)
            H[i, j] = v
            H[j, i] = np.conj(v)
    H *= O.INVCM_TO_JOULE
    lam = rng.uniform(20, 60, B) * O.INVCM_TO_JOULE
    invnu = rng.uniform(30, 120, B) * 1e-15
    om = (np.zeros(B) if omega is None else np.asarray(omega)) * (O.INVCM_TO_JOULE / O.HBAR)
    g = O.build_graph(B * K, D)
    return O.Instance(states=N, matsubaras=K, baths=B, coupling=np.asarray(coupling, np.int32),
                      hamiltonian=H, graph=g, lambda_=lam, invnu=invnu, omega=om,
                      temperature=temperature)
