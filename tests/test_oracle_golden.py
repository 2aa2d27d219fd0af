"""The oracle against its committed golden vectors (tests/golden/vectors, written by
tests/golden/make_golden.py): a change to the restatement that moves any result by more
than rounding shows up here, and the GPU tests compare the product with the same files."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from helpers import ROOT

FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "vectors", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_golden_files_exist():
    assert len(FILES) == 4


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_oracle_reproduces_golden(oracle_mod, path):
    g = load(path)
    cfg = O.parse_config(os.path.join(ROOT, str(g["config"])))
    cfg.ado_depth = int(g["depth"])
    mode = str(g["mode"])
    inst = O.instance_from_config(cfg, mode)
    if mode == "identity":
        y = inst.state_with_top(cfg.rho_init)
    else:
        _, dp = O.dipole_matrices(cfg, 0)
        ground = np.zeros((inst.states, inst.states), dtype=np.complex128)
        ground[0, 0] = 1.0
        y = inst.state_with_top(dp @ ground)
    tops = inst.propagate(y, float(g["step_size"]), int(g["steps"]), int(g["observe_steps"]))
    assert np.max(np.abs(tops - g["tops"])) <= 1e-14 * np.max(np.abs(g["tops"]))
    index = {tuple(t): u for u, t in enumerate(inst.graph.tuples)}
    s = y.reshape(inst.matrices, inst.states, inst.states, 2)
    got = np.array([s[index[tuple(t)], ..., 0] + 1j * s[index[tuple(t)], ..., 1] for t in g["tuples"]])
    assert np.max(np.abs(got - g["ados"])) <= 1e-14 * np.max(np.abs(g["ados"]))
    # the trace of the top is conserved (population runs start at trace one)
    if mode == "identity":
        assert np.max(np.abs(np.trace(g["tops"], axis1=1, axis2=2) - 1)) < 1e-12
