"""The host side of the product (config-file instance API, libheom_host.so) against the
oracle: the graph order, every kernel-input array and the filtering masks must match
bit for bit; config errors must be rejected as the reference rejects them.  No GPU."""
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import CONFIG_DIR, FMO_DIR, config_path

heom_amd = pytest.importorskip("heom_amd")

ARRAYS = ["coupling", "hamiltonian", "tuples", "plus", "minus", "pf_same_sum", "pf_same", "pf_circ", "cbk",
          "cbk_sqrt"]


def compare_instances(path, task, mode, depth=0):
    host = heom_amd.HostInstance(path, task=task, depth=depth)
    cfg = O.parse_config(path)
    if depth:
        cfg.ado_depth = depth
    ref = O.instance_from_config(cfg, mode)
    arrays = host.arrays()
    for k in ARRAYS:
        want = np.asarray(getattr(ref, k)).reshape(-1)
        assert np.array_equal(arrays[k], want), k
    assert host.matrices == ref.matrices
    assert host.info.first_non_plus == ref.first_non_plus
    return host, cfg, ref


@pytest.mark.parametrize("name,depth", [("excitonic7_d4", 0), ("excitonic8_d6", 0), ("excitonic8_d8", 5),
                                        ("excitonic24_d4", 2)])
def test_population_instance_matches_oracle(oracle_mod, name, depth):
    host, cfg, ref = compare_instances(config_path(name), heom_amd.TASK_POPULATION_DYNAMICS, "identity", depth)
    assert np.array_equal(host.top(), cfg.rho_init * 1.0)
    assert host.info.step_size == cfg.step_size and host.info.steps == cfg.steps


def test_full_size_graph_matches_oracle(oracle_mod):
    """The 735,471-ADO hierarchy (16 modes, depth 8): order, tables, prefactors."""
    compare_instances(config_path("excitonic8_d8"), heom_amd.TASK_POPULATION_DYNAMICS, "identity")


def test_fmo_examples_match_oracle(oracle_mod):
    compare_instances(FMO_DIR + "/fmo_population_dynamics_77K.cfg", heom_amd.TASK_POPULATION_DYNAMICS, "identity")
    host, cfg, ref = compare_instances(FMO_DIR + "/fmo_linear_absorption_77K.cfg",
                                       heom_amd.TASK_LINEAR_ABSORPTION, "with_ground_state")
    assert host.states == 8
    top = host.top()
    assert top[0, 0] == 1.0 and np.count_nonzero(top) == 1
    for t in range(3):
        dm, dp = host.dipoles(t)
        rdm, rdp = O.dipole_matrices(cfg, t)
        assert np.array_equal(dm, rdm) and np.array_equal(dp, rdp)


@pytest.mark.parametrize("n", range(1, 14))
@pytest.mark.parametrize("s", [1, 2, 4, 7, 16, 24])
def test_partitions_match_oracle(oracle_mod, n, s):
    assert np.array_equal(heom_amd.partitions(n, s), O.partitions(n, s))


def single_excitation_mask(graph, K, first_layer):
    """dm-heom/src/heom/hierarchy_mask.cpp:45-110 in numpy."""
    T = graph.tuples
    depth = T.sum(axis=1)
    B = T.shape[1] // K
    per_bath = T.reshape(T.shape[0], B, K).sum(axis=2)
    keep = (per_bath == depth[:, None]).any(axis=1)
    return np.where(depth < first_layer, 1, np.where(keep, 1, 0)).astype(np.uint8)


def test_masks(oracle_mod, tmp_path):
    text = open(config_path("excitonic7_d4")).read()
    assert np.all(heom_amd.HostInstance(config_path("excitonic7_d4")).mask() == 1)
    p = tmp_path / "se.cfg"
    p.write_text(text.replace("strategy=none", "strategy=single_excitation").replace("first_layer=-1", "first_layer=2"))
    host = heom_amd.HostInstance(str(p))
    g = O.build_graph(7, 4)
    assert np.array_equal(host.mask(), single_excitation_mask(g, 1, 2))


BAD = [
    ("hamiltonian=", "hamiltonian={{1, 2}, {3, 4}}\n#", "system.hamiltonien must be a system.sites"),
    ("steps=1000", "steps=1001", "remainder"),
    ("matsubaras=1", "matsubaras=0", "matsubaras"),
    ("invnu={50, 50, 50, 50, 50, 50, 50}", "invnu={50, 50, 50, 50, 50, 50, 0}", "invnu must have values > 0"),
    ("task=population_dynamics", "task=linear_absorption", "wrong program_task"),
    ("temperature=300", "temperature=300\ncolour=blue", "unrecognised option"),
    ("observe_steps=10\n", "", "required but missing"),
    ("coupling={{0}, {1}, {2}, {3}, {4}, {5}, {6}}", "coupling={{0}, {1}, {2}, {3}, {4}, {5}, {7}}",
     "baths.coupling values"),
]


@pytest.mark.parametrize("old,new,msg", BAD)
def test_config_errors(tmp_path, old, new, msg):
    text = open(config_path("excitonic7_d4")).read()
    assert old in text
    p = tmp_path / "bad.cfg"
    p.write_text(text.replace(old, new, 1))
    with pytest.raises(heom_amd.ConfigError, match=msg):
        heom_amd.HostInstance(str(p))
    with pytest.raises(O.ConfigError):
        O.parse_config(str(p))


def test_stepper_types(tmp_path):
    """solver.stepper_type takes every tableau of heom_amd.steppers (the reference's list
    comes from the un-vendored noma/num; its examples use rk_rk4) and names the accepted
    ones otherwise."""
    from heom_amd import steppers

    text = open(config_path("excitonic7_d4")).read()
    assert "stepper_type=rk_rk4" in text
    for name in steppers.TABLEAUS:
        p = tmp_path / f"{name}.cfg"
        p.write_text(text.replace("stepper_type=rk_rk4", f"stepper_type={name}", 1))
        heom_amd.HostInstance(str(p)).close()
    p = tmp_path / "bogus.cfg"
    p.write_text(text.replace("stepper_type=rk_rk4", "stepper_type=rk_bogus", 1))
    with pytest.raises(heom_amd.ConfigError, match="rk_bogus' is not supported .one of rk_euler"):
        heom_amd.HostInstance(str(p))


def test_non_hermitian_hamiltonian_rejected(tmp_path):
    text = open(config_path("excitonic7_d4")).read()
    text = re.sub(r"hamiltonian=\{\{([0-9.]+), ([0-9.-]+),", r"hamiltonian={{\1, 1.5,", text)
    p = tmp_path / "nh.cfg"
    p.write_text(text)
    with pytest.raises(heom_amd.ConfigError, match="hermitian"):
        heom_amd.HostInstance(str(p))


def test_missing_config_file():
    with pytest.raises(heom_amd.ConfigError, match="could not open"):
        heom_amd.HostInstance("/nonexistent/file.cfg")


def test_complex_entries_and_pairs(tmp_path):
    """Complex matrix entries written as (re, im) pairs parse as complex numbers."""
    text = open(config_path("excitonic7_d4")).read()
    m = re.search(r"hamiltonian=(\{.*\})", text)
    H = O.parse_literal(m.group(1))
    H[0][1] = (H[0][1], 7.25)
    H[1][0] = (H[1][0], -7.25)

    def fmt(v):
        return "(%r, %r)" % v if isinstance(v, tuple) else repr(v)

    lit = "{" + ", ".join("{" + ", ".join(fmt(v) for v in row) + "}" for row in H) + "}"
    p = tmp_path / "c.cfg"
    p.write_text(text.replace(m.group(1), lit))
    host = heom_amd.HostInstance(str(p))
    h = host.arrays()["hamiltonian"].reshape(7, 7, 2)
    assert h[0, 1, 1] == pytest.approx(7.25 * O.INVCM_TO_JOULE, rel=1e-15)
    assert h[1, 0, 1] == pytest.approx(-7.25 * O.INVCM_TO_JOULE, rel=1e-15)
    assert np.array_equal(host.arrays()["hamiltonian"], O.instance_from_config(O.parse_config(str(p))).hamiltonian)
