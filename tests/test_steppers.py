"""The explicit Runge-Kutta tableaus of heom_amd.steppers, pinned by their order of
accuracy: the HEOM operator is linear, so on a small hierarchy the exact solution is
y(T) = exp(L T) y(0), with L assembled column by column from the oracle's RHS.  Each
method's global error at T must fall as dt^p for its order p (halving dt divides it by
about 2^p), which a wrong or transposed coefficient breaks.  The GPU side
(tests/test_gpu_evaluate.py) checks that the device steppers equal these host steps."""
import numpy as np
import pytest
import scipy.linalg

from helpers import make_instance, random_state

steppers = pytest.importorskip("heom_amd.steppers")


@pytest.fixture(scope="module")
def system():
    inst = make_instance(3, 3, 1, 3, [[0], [1], [2]], seed=11)
    n = inst.state_size
    L = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[k] = 1.0
        L[:, k] = inst.rhs(e)
        e[k] = 0.0
    y0 = random_state(inst, 3)
    return inst, L, y0


@pytest.mark.parametrize("name", sorted(steppers.TABLEAUS))
def test_order_of_accuracy(system, name):
    inst, L, y0 = system
    T = 40e-15
    exact = scipy.linalg.expm(L * T) @ y0
    errs = []
    for n in (16, 32, 64):
        dt = T / n
        y = y0.copy()
        for k in range(n):
            y = steppers.numpy_step(lambda v: L @ v, y, k * dt, dt, name)
        errs.append(np.max(np.abs(y - exact)) / np.max(np.abs(exact)))
    p = steppers.TABLEAUS[name].order
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"{name}: errors {['%.2e' % e for e in errs]}, observed orders {['%.2f' % r for r in rates]}")
    assert all(abs(r - p) < 0.25 for r in rates), (name, rates)


def test_tableau_consistency():
    """Row sums of a equal c, and the b sum to 1 (every tableau is consistent)."""
    for name, t in steppers.TABLEAUS.items():
        assert abs(sum(t.b) - 1.0) < 1e-15, name
        assert t.c[0] == 0.0
        for i, row in enumerate(t.a):
            assert abs(sum(row) - t.c[i + 1]) < 1e-15, name
    assert steppers.TABLEAUS["rk_rk4"].subdiagonal and not steppers.TABLEAUS["rk_rk4_38"].subdiagonal


def test_unknown_stepper():
    with pytest.raises(ValueError, match="unknown stepper"):
        steppers.tableau("rk_does_not_exist")
