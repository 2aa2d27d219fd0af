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
    assert len(FILES) == 4  # the full-size fixture lives in tests/golden/full


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


FULL = os.path.join(ROOT, "tests", "golden", "full", "excitonic8_d8_full.npz")


def test_full_size_fixture_is_consistent(oracle_mod):
    """The full-size fixture (735,471 ADOs, 500 steps; too long to re-run here) is checked
    for what fixes it without a run: tier sizes C(d + W - 1, W - 1); the start is the
    config's rho_init; every top has trace one and is Hermitian; the tier <= 2 ADOs it
    holds are all of them, so their squared norms add up to the per-tier sums exactly;
    the norm subset starts with the top (uid 0) and never exceeds its tier's maximum."""
    from math import comb

    g = load(FULL)
    cfg = O.parse_config(os.path.join(ROOT, str(g["config"])))
    W = int(g["norm_tuples"].shape[1])
    assert W == 16 and int(g["depth"]) == cfg.ado_depth == 8
    assert list(g["tier_count"]) == [comb(d + W - 1, W - 1) for d in range(9)]
    assert sum(g["tier_count"]) == 735471
    tops = g["tops"]
    assert tops.shape == (int(g["steps"]) // int(g["observe_steps"]) + 1, 8, 8)
    assert np.array_equal(tops[0], cfg.rho_init)
    assert np.max(np.abs(np.trace(tops, axis1=1, axis2=2) - 1)) < 1e-12
    assert np.max(np.abs(tops - np.conj(np.transpose(tops, (0, 2, 1))))) < 1e-14
    tier = g["tuples"].sum(axis=1)
    nrm = np.sum(np.abs(g["ados"]) ** 2, axis=(1, 2))
    for d in range(3):
        assert int((tier == d).sum()) == g["tier_count"][d]
        assert abs(nrm[tier == d].sum() - g["tier_norm_sum"][d]) <= 1e-13 * g["tier_norm_sum"][d]
    assert not np.any(g["norm_tuples"][0]) and abs(g["norms"][0] - np.sum(np.abs(tops[-1]) ** 2)) < 1e-14
    st = g["norm_tuples"].astype(np.int64).sum(axis=1)
    assert np.all(g["norms"] <= g["tier_norm_max"][st] * (1 + 1e-15))
