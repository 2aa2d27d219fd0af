"""Time-dependent Hamiltonian on the GPU (heom_propagate_timedep): a pump/probe field
(the reference's e_field, include/heom/e_field.hpp) on the 7-site system against the
oracle's time-dependent RK4 (1e-10 relative), in both stage kernels, in a batch, and
with FP32 storage; with a zero field it is heom_propagate bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import instance, max_abs_rel

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
from heom_amd.field import gaussian_pulse, pump_probe, stage_hamiltonians  # noqa: E402


@pytest.fixture(params=["pipelined", "tile", "split"], autouse=True)
def stage_kernel(request, monkeypatch):
    monkeypatch.setenv("HEOM_STAGE_KERNEL", request.param)
    return request.param


def dipoles(N, seed=3):
    rng = np.random.default_rng(seed)
    dm = rng.normal(size=(N, N)) * 1e-21
    return dm, dm.T.copy()


FIELD = pump_probe(
    probe=dict(efield=2.0, center=40e-15, width=12e-15, phase=0.3, frequency=2.3e15),
    pump=dict(efield=5.0, center=25e-15, width=8e-15, phase=-0.5, frequency=2.25e15))


def problem(depth=3):
    cfg, inst = instance("excitonic7_d4", depth=depth)
    H0 = inst.hamiltonian.reshape(inst.states, inst.states, 2)
    H0 = H0[..., 0] + 1j * H0[..., 1]
    dm, dp = dipoles(inst.states)
    return cfg, inst, H0, dm, dp


def test_pulse_matches_oracle(oracle_mod):
    cfg, inst, H0, dm, dp = problem()
    n, dt = 120, cfg.step_size
    sh = stage_hamiltonians(H0, dm, dp, FIELD, 0.0, dt, n)
    y0 = inst.state_with_top(cfg.rho_init)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y0)
        tops = h.propagate_timedep(dt, n, 10, sh)
        h.set_state(y0)
        plain = h.propagate(dt, n, 10)
    ref = inst.propagate_timedep(y0.copy(), dt, n, 10, sh)
    err = max_abs_rel(tops, ref)
    print(f"pump/probe field, 120 steps: max rel deviation from the oracle {err:.2e}; "
          f"field effect {max_abs_rel(plain, ref):.2e}", flush=True)
    assert err < 1e-10
    assert max_abs_rel(plain, ref) > 1e-6  # the field changes the dynamics


def test_zero_field_is_propagate(oracle_mod):
    cfg, inst, H0, dm, dp = problem()
    sh = stage_hamiltonians(H0, dm, dp, lambda t: 0.0 * t, 0.0, cfg.step_size, 40)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(inst.state_with_top(cfg.rho_init))
        a = h.propagate_timedep(cfg.step_size, 40, 10, sh)
        h.set_state(inst.state_with_top(cfg.rho_init))
        b = h.propagate(cfg.step_size, 40, 10)
    assert np.array_equal(a, b)


def test_batch_and_fp32(oracle_mod):
    cfg, a, H0, dm, dp = problem(depth=3)
    _, b, _, _, _ = problem(depth=2)
    n, dt = 60, cfg.step_size
    sha = stage_hamiltonians(H0, dm, dp, FIELD, 0.0, dt, n)
    shb = stage_hamiltonians(H0 * 1.0001, dm, dp, FIELD, 5e-15, dt, n)
    with heom_amd.Handle.from_instances([a, b]) as hb:
        hb.clear_state()
        hb.set_top(np.array([cfg.rho_init, cfg.rho_init]))
        rec = hb.propagate_timedep(dt, n, 20, np.stack([sha, shb], axis=2))
    for k, (inst, sh) in enumerate(((a, sha), (b, shb))):
        ref = inst.propagate_timedep(inst.state_with_top(cfg.rho_init), dt, n, 20, sh)
        assert max_abs_rel(rec[:, k], ref) < 1e-10, k
    with heom_amd.Handle.from_instance(a, precision=heom_amd.HEOM_PRECISION_FP32) as h:
        h.set_state(a.state_with_top(cfg.rho_init))
        t32 = h.propagate_timedep(dt, n, 20, sha)
    ref = a.propagate_timedep(a.state_with_top(cfg.rho_init), dt, n, 20, sha)
    perr = np.max(np.abs(np.real(np.diagonal(t32 - ref, axis1=1, axis2=2))))
    print(f"FP32 storage with the field: population error {perr:.2e}", flush=True)
    assert perr < 1e-6
