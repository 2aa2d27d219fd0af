"""The oracle's time-dependent Hamiltonian (oracle_rk4_step_timedep,
oracle_propagate_timedep) pinned on the CPU: with a constant field it is the plain RK4
step bit for bit, and with a smooth pulse it converges at fourth order, which it does only
if every stage sees the Hamiltonian at its own time t, t + h/2, t + h/2, t + h."""
import numpy as np
import pytest

import oracle as O
from test_oracle import small_instance

heom_amd = pytest.importorskip("heom_amd")
from heom_amd.field import gaussian_pulse, stage_hamiltonians  # noqa: E402


def system(seed=2):
    inst, H = small_instance(3, 1, 2, [[0], [1], [2]], seed=seed)
    rng = np.random.default_rng(seed)
    dm = rng.normal(size=(3, 3)) * 2e-21
    dp = dm.T.copy()
    field = gaussian_pulse(efield=3.0, center=30e-15, width=10e-15, phase=0.4, frequency=2.2e15)
    return inst, H, dm, dp, field


def test_constant_field_equals_plain_rk4(oracle_mod):
    inst, H, dm, dp, _ = system()
    y = np.random.default_rng(1).normal(size=inst.state_size)
    sh = stage_hamiltonians(H, dm, dp, lambda t: 0.0 * t, 0.0, 0.5e-15, 40)
    a = inst.propagate_timedep(y.copy(), 0.5e-15, 40, 10, sh)
    b = inst.propagate(y.copy(), 0.5e-15, 40, 10)
    assert np.array_equal(a, b)


def test_pulse_converges_at_fourth_order(oracle_mod):
    inst, H, dm, dp, field = system()
    y0 = inst.state_with_top(np.diag([0.2, 0.5, 0.3]).astype(np.complex128))
    T = 60e-15

    def run(n):
        h = T / n
        return inst.propagate_timedep(y0.copy(), h, n, n, stage_hamiltonians(H, dm, dp, field, 0.0, h, n))[-1]

    ref = run(7680)
    errs = [np.max(np.abs(run(n) - ref)) for n in (240, 480, 960)]
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    print("errors", errs, "ratios", ratios)
    assert all(12 < r < 20 for r in ratios), ratios
    # the field matters: a run with E = 0 ends elsewhere
    flat = inst.propagate_timedep(y0.copy(), T / 960, 960, 960, stage_hamiltonians(H, dm, dp, lambda t: 0 * t, 0.0,
                                                                                    T / 960, 960))[-1]
    assert np.max(np.abs(flat - ref)) > 100 * errs[-1]


def test_field_helpers_match_the_oracle_restatement(oracle_mod):
    inst, H, dm, dp, field = system()
    a = stage_hamiltonians(H, dm, dp, field, 1e-15, 0.7e-15, 9)
    b = O.field_stage_hamiltonians(H, dm, dp, field, 1e-15, 0.7e-15, 9)
    assert np.allclose(a, b, rtol=1e-15, atol=0)
    # Hermitian when D+ = D-^T (real dipoles)
    assert np.allclose(a, np.conj(np.swapaxes(a, -1, -2)), rtol=0, atol=1e-12 * np.abs(H).max())
