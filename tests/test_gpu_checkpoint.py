"""Checkpoint and resume through the reference's binary instance file: 50 RK4 steps on the
GPU, the whole hierarchy written by the host library and read back into a fresh handle,
50 more steps -- bitwise the state and tops of one 100-step run (the file holds the FP64
state exactly and the step sequence does not depend on where a run starts)."""
import numpy as np
import pytest

from helpers import config_path

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


@pytest.mark.parametrize("name,depth", [("excitonic8_d6", 4), ("excitonic24_d4", 2)])
def test_checkpoint_resume_is_bitwise(tmp_path, name, depth):
    host = heom_amd.HostInstance(config_path(name), depth=depth)
    N, M = host.states, host.matrices
    dt = float(host.info.step_size)
    y0 = np.zeros(M * N * N * 2)
    top = host.top()
    y0[:N * N * 2] = np.stack([top.real, top.imag], -1).reshape(-1)
    with host.create_handle() as h:
        h.set_state(y0)
        full = h.propagate(dt, 100, 50)
        y100 = h.get_state()
    f = tmp_path / "checkpoint.bin"
    with host.create_handle() as h:
        h.set_state(y0)
        first = h.propagate(dt, 50, 50)
        host.write_instance(f, h.get_state(), time=50 * dt)
    state, t = host.read_instance_state(f)
    assert t == 50 * dt
    with host.create_handle() as h:
        h.set_state(state)
        rest = h.propagate(dt, 50, 50)
        y_resumed = h.get_state()
    print(f"{name} depth {depth} ({M} ADOs): resume after 50 steps, max |diff| "
          f"{np.max(np.abs(y_resumed - y100)):.1e}", flush=True)
    assert np.array_equal(first, full[:2])
    assert np.array_equal(rest, full[1:])
    assert np.array_equal(y_resumed, y100)
