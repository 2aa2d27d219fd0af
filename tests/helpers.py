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
