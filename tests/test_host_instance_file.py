"""The reference's binary instance file (src/heom/instance.cpp:281-377) written and read
by the host library (no GPU), against the oracle's own reader and writer of the format:
the header fields and sizes, every kernel-input array, the hierarchy (whole or top
only), and the checks that refuse a file of another instance or a truncated one."""
import numpy as np
import pytest

import oracle as O
from helpers import FMO_DIR, config_path, instance, random_state

heom_amd = pytest.importorskip("heom_amd")

CASES = [(config_path("excitonic8_d6"), heom_amd.TASK_POPULATION_DYNAMICS, 3, "identity"),
         (FMO_DIR + "/fmo_linear_absorption_77K.cfg", heom_amd.TASK_LINEAR_ABSORPTION, 2, "with_ground_state")]


def test_header_layout_is_packed():
    # version, top_only (bool), two layouts and the states mode, time, first_non_plus_id,
    # then seven int64 sizes, written field by field with no padding
    assert O._INSTANCE_HEADER.itemsize == 4 + 1 + 4 + 4 + 4 + 8 + 4 + 7 * 8


@pytest.mark.parametrize("path,task,depth,mode", CASES)
@pytest.mark.parametrize("top_only", [False, True])
def test_host_file_read_by_oracle(oracle_mod, tmp_path, path, task, depth, mode, top_only):
    host = heom_amd.HostInstance(path, task=task, depth=depth)
    N, M = host.states, host.matrices
    rng = np.random.default_rng(5)
    state = rng.normal(size=M * N * N * 2)
    f = tmp_path / "inst.bin"
    host.write_instance(f, state, top_only=top_only, time=1.25e-13)
    got = O.read_instance_file(str(f))
    d = host.desc()
    assert f.stat().st_size == O.instance_file_size(N, M, d.matsubaras, d.baths, d.width, top_only)
    assert got["version"] == 0 and got["hierarchy_layout"] == 0 and got["hamiltonian_layout"] == 0
    assert got["top_only"] == top_only
    assert got["sts_mode"] == O.STS_MODES[mode]
    assert got["time"] == 1.25e-13
    assert (got["states"], got["matrices"], got["matsubaras"], got["baths"], got["width"], got["depth"]) == \
        (N, M, d.matsubaras, d.baths, d.width, depth)
    assert got["baths_per_state"] == d.baths_max_per_state
    assert np.array_equal(got["hierarchy"], state[:N * N * 2] if top_only else state)
    arrays = host.arrays()
    for k in ("hamiltonian", "tuples", "plus", "minus", "pf_same_sum", "pf_same", "pf_circ", "cbk", "cbk_sqrt"):
        assert np.array_equal(got[k], arrays[k]), k
    # the tables are the oracle's graph for the same width and depth
    g = O.build_graph(d.width, depth)
    assert np.array_equal(got["tuples"], g.tuples.reshape(-1))
    assert np.array_equal(got["plus"], g.plus.reshape(-1))
    assert np.array_equal(got["minus"], g.minus.reshape(-1))
    assert got["first_non_plus"] == g.first_non_plus


def test_oracle_file_read_by_host(oracle_mod, tmp_path):
    _, inst = instance("excitonic8_d6", depth=3)
    host = heom_amd.HostInstance(config_path("excitonic8_d6"), depth=3)
    y = random_state(inst, 11)
    f = tmp_path / "oracle.bin"
    O.write_instance_file(str(f), inst, y, time=2.5e-14)
    state, t = host.read_instance_state(f)
    assert t == 2.5e-14
    assert np.array_equal(state, y)
    O.write_instance_file(str(f), inst, y, top_only=True, time=0.0)
    state, t = host.read_instance_state(f)
    n2 = inst.states * inst.states * 2
    assert np.array_equal(state[:n2], y[:n2]) and not state[n2:].any()


def test_host_round_trip_and_top_default(tmp_path):
    host = heom_amd.HostInstance(config_path("excitonic8_d6"), depth=2)
    N, M = host.states, host.matrices
    y = np.random.default_rng(2).normal(size=M * N * N * 2)
    f = tmp_path / "a.bin"
    host.write_instance(f, y, time=3.0)
    state, t = host.read_instance_state(f)
    assert t == 3.0 and np.array_equal(state, y)
    host.write_instance(f)  # no state: the config's initial top ADO
    state, t = host.read_instance_state(f)
    top = host.top()
    assert np.array_equal(state[:N * N * 2], np.stack([top.real, top.imag], -1).reshape(-1))
    assert not state[N * N * 2:].any() and t == 0.0


def test_mismatched_or_truncated_files_are_refused(tmp_path):
    a = heom_amd.HostInstance(config_path("excitonic8_d6"), depth=2)
    b = heom_amd.HostInstance(config_path("excitonic8_d6"), depth=3)
    f = tmp_path / "a.bin"
    a.write_instance(f, np.zeros(a.matrices * a.states * a.states * 2))
    with pytest.raises(heom_amd.ConfigError, match="does not match"):
        b.read_instance_state(f)
    data = f.read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[:-8])
    with pytest.raises(heom_amd.ConfigError, match="truncated"):
        a.read_instance_state(tmp_path / "cut.bin")
    (tmp_path / "long.bin").write_bytes(data + b"\0")
    with pytest.raises(heom_amd.ConfigError, match="trailing"):
        a.read_instance_state(tmp_path / "long.bin")
    with pytest.raises(heom_amd.ConfigError, match="could not open"):
        a.read_instance_state(tmp_path / "missing.bin")
    with pytest.raises(ValueError):
        a.write_instance(f, np.zeros(5))
