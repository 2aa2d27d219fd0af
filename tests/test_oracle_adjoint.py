"""The oracle's adjoint (heom_oracle.c: oracle_rhs_adjoint, oracle_adjoint_contract,
oracle_adjoint_gradient) pinned on the CPU:

* L^dag against the conjugate transpose of the dense operator-form Liouvillian that
  already pins the forward RHS, and <mu, L y> = <L^dag mu, y> with a filtering mask;
* every coefficient gradient of <mu, L y> against central differences of the RHS;
* the discrete-adjoint gradient of J = sum_r Re(w_r Tr(O sigma_0(t_r))) against central
  differences of whole RK4 trajectories, for every coefficient array, the initial
  state and the physical bath parameters (lambda, gamma, Omega).
"""
import numpy as np
import pytest

import oracle as O
from test_oracle import dense_liouvillian, small_instance, to_dense

CASES = [
    dict(N=3, K=1, D=3, coupling=[[0], [1], [2]]),
    dict(N=2, K=2, D=3, coupling=[[0], [1]]),
    dict(N=3, K=2, D=2, coupling=[[0], [1], [2]], omega=[0, 120, -80], complex_h=True),
    dict(N=3, K=1, D=2, coupling=[[-1], [0], [1]]),
    dict(N=3, K=1, D=2, coupling=[[0, 1], [1, -1], [0, -1]]),
]


def inner(a, b):
    """<a, b> = Re sum conj(a) b on interleaved states."""
    return float(np.dot(a, b))


def make(case, seed=3):
    return small_instance(case["N"], case["K"], case["D"], case["coupling"], seed=seed,
                          omega=case.get("omega"), complex_h=case.get("complex_h", False))


@pytest.mark.parametrize("case", CASES)
def test_rhs_adjoint_is_the_conjugate_transpose(oracle_mod, case):
    inst, H = make(case)
    L, ados, _ = dense_liouvillian(inst, H)
    mu = np.random.default_rng(4).normal(size=inst.state_size)
    got = to_dense(inst, inst.rhs_adjoint(mu), ados)
    want = L.conj().T @ to_dense(inst, mu, ados)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_rhs_adjoint_identity_with_mask(oracle_mod):
    inst, _ = make(dict(N=3, K=2, D=3, coupling=[[0], [1], [2]]))
    rng = np.random.default_rng(7)
    inst.mask[:] = (rng.random(inst.matrices) > 0.3).astype(np.uint8)
    inst.mask[0] = O.MASK_TRUE
    mu, y = rng.normal(size=(2, inst.state_size))
    a, b = inner(mu, inst.rhs(y)), inner(inst.rhs_adjoint(mu), y)
    assert abs(a - b) <= 1e-13 * np.linalg.norm(mu) * np.linalg.norm(inst.rhs(y))


ARRAYS = ("hamiltonian", "pf_same_sum", "pf_same", "pf_circ", "cbk")


def directional(inst, key, grad, direction):
    """Re sum conj(G) dk for a real direction over the interleaved array."""
    d = direction[0::2] + 1j * direction[1::2]
    return float(np.sum(np.real(np.conj(np.ravel(grad)) * d)))


@pytest.mark.parametrize("case", CASES)
def test_contract_matches_central_differences_of_the_rhs(oracle_mod, case):
    inst, _ = make(case)
    rng = np.random.default_rng(11)
    mu, y = rng.normal(size=(2, inst.state_size))
    g = inst.adjoint_contract(mu, y, 0.7)
    for key in ARRAYS:
        arr = getattr(inst, key)
        d = rng.normal(size=arr.size) * np.abs(arr).max()
        # <mu, L y> is linear in every array but cbk: a large step is exact there and
        # keeps the commutator's magnitude from cancelling the difference away
        eps = 1e-6 if key == "cbk" else 0.25
        base = arr.copy()
        arr[:] = base + eps * d
        fp = inner(mu, inst.rhs(y))
        arr[:] = base - eps * d
        fm = inner(mu, inst.rhs(y))
        arr[:] = base
        fd = 0.7 * (fp - fm) / (2 * eps)
        got = directional(inst, key, g[key], d)
        scale = abs(fd) + 1e-30
        assert abs(got - fd) <= 1e-6 * scale, (key, got, fd)


def objective(inst, y0, h, nsteps, obs, op, w):
    tops = inst.propagate(y0.copy(), h, nsteps, obs)
    X = np.array([np.trace(op @ t) for t in tops])
    return float(np.sum(np.real(w * X))), X


def gradient_case():
    inst, _ = make(dict(N=3, K=2, D=3, coupling=[[0], [1], [2]], omega=[0, 60, 0]), seed=8)
    rng = np.random.default_rng(21)
    N = inst.states
    y0 = np.zeros(inst.state_size)
    rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = rho @ rho.conj().T
    rho /= np.trace(rho)
    y0[:2 * N * N] = np.stack([rho.real, rho.imag], -1).reshape(-1)
    op = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    nsteps, obs = 37, 3
    rows = nsteps // obs + 1
    w = rng.normal(size=rows) + 1j * rng.normal(size=rows)
    return inst, y0, 1e-15, nsteps, obs, op, w


def test_adjoint_gradient_matches_trajectory_differences(oracle_mod):
    inst, y0, h, nsteps, obs, op, w = gradient_case()
    X, g = inst.adjoint_gradient(y0, h, nsteps, obs, op, w, checkpoint_steps=5)
    J0, X0 = objective(inst, y0, h, nsteps, obs, op, w)
    assert np.max(np.abs(X - X0)) <= 1e-14 * np.max(np.abs(X0))
    rng = np.random.default_rng(5)
    for key in ARRAYS:
        arr = getattr(inst, key)
        d = rng.normal(size=arr.size) * np.abs(arr).max()
        eps = 1e-6
        base = arr.copy()
        arr[:] = base + eps * d
        jp, _ = objective(inst, y0, h, nsteps, obs, op, w)
        arr[:] = base - eps * d
        jm, _ = objective(inst, y0, h, nsteps, obs, op, w)
        arr[:] = base
        fd = (jp - jm) / (2 * eps)
        got = directional(inst, key, g[key], d)
        print(f"{key}: adjoint {got:.10e} differences {fd:.10e}")
        assert abs(got - fd) <= 1e-6 * abs(fd), key
    # initial state
    d = rng.normal(size=inst.state_size) * (0.3 ** inst.graph.tuples.sum(axis=1)).repeat(2 * inst.states ** 2)
    jp, _ = objective(inst, y0 + 1e-3 * d, h, nsteps, obs, op, w)
    jm, _ = objective(inst, y0 - 1e-3 * d, h, nsteps, obs, op, w)
    fd = (jp - jm) / 2e-3
    assert abs(inner(g["state"], d) - fd) <= 1e-9 * abs(fd)


@pytest.mark.parametrize("ck", [1, 4, 13, 100])
def test_checkpoint_interval_does_not_change_the_gradient(oracle_mod, ck):
    inst, y0, h, nsteps, obs, op, w = gradient_case()
    _, ref = inst.adjoint_gradient(y0, h, nsteps, obs, op, w, checkpoint_steps=50)
    _, g = inst.adjoint_gradient(y0, h, nsteps, obs, op, w, checkpoint_steps=ck)
    for key in ARRAYS + ("state",):
        assert np.allclose(g[key], ref[key], rtol=1e-12, atol=1e-12 * np.max(np.abs(ref[key]))), key


def test_bath_parameter_gradient_matches_trajectory_differences(oracle_mod):
    inst, y0, h, nsteps, obs, op, w = gradient_case()
    _, g = inst.adjoint_gradient(y0, h, nsteps, obs, op, w, checkpoint_steps=10)
    got = inst.bath_parameter_gradient(g)
    base = np.stack([inst.lambda_ / O.INVCM_TO_JOULE, O.HBAR / (inst.invnu * O.INVCM_TO_JOULE),
                     inst.omega * O.HBAR / O.INVCM_TO_JOULE])
    Hm = inst.hamiltonian.reshape(inst.states, inst.states, 2)
    Hc = Hm[..., 0] + 1j * Hm[..., 1]

    def J(p):
        other = O.Instance(states=inst.states, matsubaras=inst.matsubaras, baths=inst.baths,
                           coupling=inst.coupling, hamiltonian=Hc, graph=inst.graph,
                           lambda_=p[0] * O.INVCM_TO_JOULE, invnu=O.HBAR / (p[1] * O.INVCM_TO_JOULE),
                           omega=p[2] * O.INVCM_TO_JOULE / O.HBAR, temperature=inst.temperature)
        return objective(other, y0, h, nsteps, obs, op, w)[0]

    for k in range(3):
        for b in range(inst.baths):
            step = 1e-3
            p = base.copy()
            p[k, b] += step
            jp = J(p)
            p[k, b] -= 2 * step
            jm = J(p)
            fd = (jp - jm) / (2 * step)
            print(f"param {k} bath {b}: adjoint {got[k, b]:.10e} differences {fd:.10e}")
            assert abs(got[k, b] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(got))), (k, b)
