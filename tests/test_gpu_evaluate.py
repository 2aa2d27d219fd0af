"""The per-evaluation interface (heom_evaluate, heom_lincomb, heom_swap_buffers,
heom_add_buffers, heom_get_buffer / heom_set_buffer, the pre-evaluate actions): the
reference's ode::solve forms (dm-heom/src/heom/ode.cpp:97-196, heom_ode.cl:416-434) on
device buffers, and explicit Runge-Kutta steppers built on them (heom_amd.steppers).

* every form of heom_evaluate against the oracle's RHS (1e-13 relative), on the tile or
  pipelined kernel (7 states) and on the large-system kernel (9 states);
* heom_lincomb against numpy;
* classical RK4 composed from heom_evaluate against the fused heom_step_rk4 (1e-13);
* every tableau on the device against the same tableau on the host with the oracle's RHS
  (1e-12 over 10 steps);
* a time-dependent Hamiltonian set by a pre-evaluate action against
  heom_propagate_timedep with the same stage Hamiltonians (1e-12), and the action's
  arguments;
* the FP32 and argument errors."""
import numpy as np
import pytest

import oracle as O
from helpers import instance, make_instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
from heom_amd import steppers  # noqa: E402
from heom_amd.field import gaussian_pulse, stage_hamiltonians  # noqa: E402

DT = 1e-15


def cases():
    yield "7 states", instance("excitonic7_d4", depth=3)[1]
    yield "9 states", make_instance(9, 9, 1, 2, [[i] for i in range(9)], seed=4)


@pytest.fixture(params=list(range(2)), ids=["7 states", "9 states"])
def inst(request):
    return list(cases())[request.param][1]


def test_evaluate_forms(oracle_mod, inst):
    y = random_state(inst, 1)
    base = random_state(inst, 2)
    acc0 = random_state(inst, 3)
    f = inst.rhs(y)
    with heom_amd.Handle.from_instance(inst) as h:
        assert h.buffer_count() == 4
        h.set_state(y)
        h.set_buffer(2, base)
        h.set_buffer(3, acc0)
        # pure evaluation (ode::solve(t, h, in, out)) equals heom_rhs bit for bit
        h.evaluate(0.0, DT, 0, out=1)
        pure = h.get_buffer(1)
        assert np.array_equal(pure, h.rhs())
        assert max_abs_rel(pure, f) < 1e-13
        # out = base + c f with acc = acc + c' f (the sub-diagonal and accumulating forms)
        h.evaluate(0.0, DT, 0, out=1, out_base=2, out_coeff=0.7 * DT, acc=3, acc_coeff=-0.3 * DT)
        assert max_abs_rel(h.get_buffer(1), base + 0.7 * DT * f) < 1e-13
        assert max_abs_rel(h.get_buffer(3), acc0 - 0.3 * DT * f) < 1e-13
        # acc_init: acc = in + c f; out without base: out = c f; the state is unchanged
        h.evaluate(0.0, DT, 0, out=2, out_coeff=2.0, acc=1, acc_coeff=DT, acc_init=True)
        assert max_abs_rel(h.get_buffer(2), 2.0 * f) < 1e-13
        assert max_abs_rel(h.get_buffer(1), y + DT * f) < 1e-13
        assert np.array_equal(h.get_state(), y)
        # in-place: out = in (the stage input is consumed)
        h.set_buffer(1, y)
        h.evaluate(0.0, DT, 1, out=1, out_base=1, out_coeff=DT)
        assert max_abs_rel(h.get_buffer(1), y + DT * f) < 1e-13


def test_lincomb_and_buffers(oracle_mod):
    _, inst = instance("excitonic7_d4", depth=2)
    ys = [random_state(inst, s) for s in range(4)]
    with heom_amd.Handle.from_instance(inst) as h:
        assert h.add_buffers(3) == 7 and h.buffer_count() == 7
        for b in range(4):
            h.set_buffer(b + 3, ys[b])
        h.lincomb(0, [(0.5, 3), (-2.0, 4), (1.25, 5), (3.0, 6)])
        want = 0.5 * ys[0] - 2.0 * ys[1] + 1.25 * ys[2] + 3.0 * ys[3]
        assert max_abs_rel(h.get_state(), want) < 1e-15
        h.lincomb(3, [(2.0, 3)])  # dst among the sources
        assert max_abs_rel(h.get_buffer(3), 2.0 * ys[0]) < 1e-16
        h.swap_buffers(0, 4)
        assert np.array_equal(h.get_state(), ys[1]) and max_abs_rel(h.get_buffer(4), want) < 1e-15
        with pytest.raises(heom_amd.HeomError, match="no state buffer"):
            h.lincomb(7, [(1.0, 0)])
        with pytest.raises(heom_amd.HeomError, match="1 to 8"):
            h.lincomb(0, [(1.0, 0)] * 9)
        with pytest.raises(heom_amd.HeomError, match="different"):
            h.evaluate(0.0, DT, 0, out=1, acc=1, acc_coeff=1.0)
        with pytest.raises(heom_amd.HeomError, match="at most 16"):
            h.add_buffers(10)


def test_rk4_composed_matches_fused(oracle_mod, inst):
    y = random_state(inst, 5)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        h.step_rk4(DT, 20)
        fused = h.get_state()
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        rk = steppers.ExplicitRK(h, "rk_rk4")
        for n in range(20):
            rk.step(n * DT, DT)
        composed = h.get_state()
    err = max_abs_rel(composed, fused)
    print(f"RK4 composed vs fused, 20 steps: {err:.1e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("name", sorted(steppers.TABLEAUS))
def test_tableau_on_device(oracle_mod, name):
    _, inst = instance("excitonic7_d4", depth=3)
    y = random_state(inst, 6)
    ref = y.copy()
    for n in range(10):
        ref = steppers.numpy_step(inst.rhs, ref, n * DT, DT, name)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        rk = steppers.ExplicitRK(h, name)
        for n in range(10):
            rk.step(n * DT, DT)
        got = h.get_state()
    err = max_abs_rel(got, ref)
    print(f"{name}: 10 steps against the host tableau {err:.1e}", flush=True)
    assert err < 1e-12


def test_pre_evaluate_time_dependent_hamiltonian(oracle_mod):
    """A field applied by a pre-evaluate action (the reference's e_field action,
    dm-heom/include/heom/e_field.hpp:63-87) equals heom_propagate_timedep."""
    _, inst = instance("excitonic7_d4", depth=3)
    N = inst.states
    H0 = inst.hamiltonian_matrix()
    rng = np.random.default_rng(8)
    dm = np.triu(rng.normal(size=(N, N)), 1) * 0.02 * O.INVCM_TO_JOULE * 1e3
    dp = dm.conj().T
    field = gaussian_pulse(efield=1.0, center=10e-15, width=4e-15, phase=0.3, frequency=2e14)
    steps, obs = 30, 5
    stage_h = stage_hamiltonians(H0, dm, dp, field, 0.0, DT, steps)
    top = np.zeros((N, N), complex)
    top[0, 0] = 1.0
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_top(top)
        want = h.propagate_timedep(DT, steps, obs, stage_h)
    calls = []
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_top(top)

        def action(handle, time, step_size, src, out):
            calls.append((time, step_size, src, out))
            E = complex(field(time))
            handle.set_hamiltonian(H0 - E * dm - np.conj(E) * dp)

        h.add_pre_evaluate(action)
        got = steppers.ExplicitRK(h, "rk_rk4").propagate(0.0, DT, steps, obs)
        h.clear_pre_evaluate()
        h.evaluate(0.0, DT, 0, out=1)  # no action any more
    assert len(calls) == 4 * steps
    times = np.array([c[0] for c in calls]).reshape(steps, 4)
    assert np.allclose(times, DT * (np.arange(steps)[:, None] + np.array([0, 0.5, 0.5, 1.0])[None]), rtol=1e-13,
                       atol=0)
    assert all(c[1] == DT for c in calls) and [c[2] for c in calls[:4]] == [0, 2, 3, 2]
    assert got.shape == want.shape
    err = max_abs_rel(got, want)
    print(f"field through pre-evaluate actions vs heom_propagate_timedep: {err:.1e}", flush=True)
    assert err < 1e-12


def test_pre_evaluate_failure_and_fp32(oracle_mod):
    _, inst = instance("excitonic7_d4", depth=2)
    with heom_amd.Handle.from_instance(inst) as h:
        h.add_pre_evaluate(lambda *a: 3)
        with pytest.raises(heom_amd.HeomError, match="pre-evaluate action returned 3"):
            h.evaluate(0.0, DT, 0, out=1)
    with heom_amd.Handle.from_instance(inst, precision=heom_amd.HEOM_PRECISION_FP32) as h:
        with pytest.raises(heom_amd.HeomError, match="FP64 handles"):
            h.evaluate(0.0, DT, 0, out=1)
