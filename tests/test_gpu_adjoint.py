"""The adjoint on the GPU (heom_rhs_adjoint, heom_adjoint_forward/backward and the torch
wrapper heom_amd.adjoint) against the oracle's adjoint (FP64 reference gradient) and
against central finite differences of GPU trajectories.

The adjoint-gradient workload: 8 sites, 8 baths with one mode each, depth 6 (3003
ADOs), RK4 dt = 1 fs for 1000 steps, checkpoints every 50 steps; J = the time integral
of the population of site 3 over 1 ps; gradients with respect to the 8 site energies,
the 28 couplings and the 8 (lambda, gamma) pairs.  The bar is 1e-6 relative, for the
finite differences with h = 1e-3 cm^-1 and for the oracle's gradient.
"""
import numpy as np
import pytest

import oracle as O
from helpers import config_path, instance, max_abs_rel, random_state
from test_oracle import small_instance

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
torch = pytest.importorskip("torch")
from heom_amd.adjoint import System  # noqa: E402


def inner(a, b):
    return float(np.dot(a, b))


ARRAYS = ("pf_same_sum", "pf_same", "pf_circ", "cbk")

SMALL = [
    dict(N=3, K=2, D=3, coupling=[[0], [1], [2]], omega=[0, 60, 0]),
    dict(N=3, K=2, D=2, coupling=[[0], [1], [2]], omega=[0, 120, -80], complex_h=True),
    dict(N=3, K=1, D=2, coupling=[[-1], [0], [1]]),
    dict(N=3, K=1, D=2, coupling=[[0, 1], [1, -1], [0, -1]]),
    # eight sites with two Matsubara terms per bath: four terms per element at N = 8
    dict(N=8, K=2, D=2, coupling=[[i] for i in range(8)]),
]


def small(case, seed=3):
    return small_instance(case["N"], case["K"], case["D"], case["coupling"], seed=seed, omega=case.get("omega"),
                          complex_h=case.get("complex_h", False))[0]


@pytest.mark.parametrize("which", ["excitonic8_d6", "excitonic8_d8"] + [f"small{k}" for k in range(len(SMALL))])
def test_rhs_adjoint_matches_oracle(oracle_mod, which):
    if which.startswith("small"):
        inst = small(SMALL[int(which[5:])])
    else:
        _, inst = instance(which, depth=3 if which == "excitonic8_d8" else None)
    mu = random_state(inst, 5)
    ref = inst.rhs_adjoint(mu)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(mu)
        got = h.rhs_adjoint()
        sigma = random_state(inst, 6)
        h.set_state(sigma)
        f = h.rhs()
    err = max_abs_rel(got, ref)
    print(f"{which}: L^dag max rel deviation from the oracle {err:.2e}", flush=True)
    assert err < 1e-12
    # <mu, L sigma> = <L^dag mu, sigma> with both operators from the GPU
    a, b = inner(mu, f), inner(got, sigma)
    assert abs(a - b) <= 1e-12 * np.linalg.norm(mu) * np.linalg.norm(f)


def gradient_problem(inst, seed=21, nsteps=37, obs=3):
    rng = np.random.default_rng(seed)
    N = inst.states
    rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = rho @ rho.conj().T
    rho /= np.trace(rho)
    y0 = inst.state_with_top(rho)
    op = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rows = nsteps // obs + 1
    w = rng.normal(size=rows) + 1j * rng.normal(size=rows)
    return y0, op, w, nsteps, obs


def check_against_oracle(X, g, Xr, gr, tol=1e-10):
    assert max_abs_rel(X, Xr) < 1e-12
    errs = {"hamiltonian": max_abs_rel(g["hamiltonian"], gr["hamiltonian"])}
    for k in ARRAYS:
        errs[k] = max_abs_rel(g[k], gr[k])
    errs["state"] = max_abs_rel(g["state"], gr["state"])
    print("gradient deviations from the oracle: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < tol, k


# the forward sweep keeps every step's stage vectors when they fit on the device (keep);
# HEOM_ADJOINT_RECOMPUTE=1 selects checkpoints and recomputation instead
@pytest.mark.parametrize("mode", ["keep", "recompute"])
@pytest.mark.parametrize("k", range(len(SMALL)))
@pytest.mark.parametrize("ck", [5, 100])
def test_gradients_match_oracle(oracle_mod, monkeypatch, k, ck, mode):
    if mode == "recompute":
        monkeypatch.setenv("HEOM_ADJOINT_RECOMPUTE", "1")
    inst = small(SMALL[k], seed=8)
    y0, op, w, nsteps, obs = gradient_problem(inst)
    Xr, gr = inst.adjoint_gradient(y0, 1e-15, nsteps, obs, op, w, checkpoint_steps=ck)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y0)
        X = h.adjoint_forward(1e-15, nsteps, obs, op, checkpoint_steps=ck)[:, 0]
        end = h.get_state()
        g = h.adjoint_backward(w)
        # a second backward over the same record, with other weights, is independent
        g2 = h.adjoint_backward(2.0 * w)
    check_against_oracle(X, {**g, "hamiltonian": g["hamiltonian"][0]}, Xr, gr)
    assert np.allclose(g2["cbk"], 2.0 * g["cbk"], rtol=1e-12, atol=0)
    y = y0.copy()
    inst.propagate(y, 1e-15, nsteps, obs)
    assert max_abs_rel(end, y) < 1e-12


@pytest.mark.parametrize("ck", [1, 6, 50])
def test_trajectory_and_recompute_modes_agree_bitwise(oracle_mod, monkeypatch, ck):
    """Both modes run the same kernels on the same inputs: the gradients, observations and
    final states must be identical, for checkpoint intervals that do and do not divide the
    step count and observation interval."""
    inst = small(SMALL[1], seed=5)
    y0, op, w, nsteps, obs = gradient_problem(inst, seed=4, nsteps=41, obs=4)
    out = {}
    for mode in ("keep", "recompute"):
        if mode == "recompute":
            monkeypatch.setenv("HEOM_ADJOINT_RECOMPUTE", "1")
        with heom_amd.Handle.from_instance(inst) as h:
            h.set_state(y0)
            X = h.adjoint_forward(1e-15, nsteps, obs, op, checkpoint_steps=ck)
            end = h.get_state()
            g = h.adjoint_backward(w)
            g2 = h.adjoint_backward(w[::-1].copy())
        out[mode] = (X, end, g, g2)
    a, b = out["keep"], out["recompute"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for ga, gb in ((a[2], b[2]), (a[3], b[3])):
        for key in ("hamiltonian", "state") + ARRAYS:
            assert np.array_equal(ga[key], gb[key]), key


def test_batch_gradients_match_oracle_per_instance(oracle_mod):
    insts = [small(SMALL[0], seed=s) for s in (8, 9, 10)]
    # same baths, different Hamiltonians (small_instance draws the baths from the seed too)
    for inst in insts[1:]:
        for key in ("pf_same_sum", "pf_same", "pf_circ", "cbk"):
            getattr(inst, key)[:] = getattr(insts[0], key)
    probs = [gradient_problem(inst, seed=30 + i) for i, inst in enumerate(insts)]
    nsteps, obs = probs[0][3], probs[0][4]
    with heom_amd.Handle.from_instances(insts) as h:
        h.set_state(np.concatenate([p[0] for p in probs]))
        X = h.adjoint_forward(1e-15, nsteps, obs, np.array([p[1] for p in probs]), checkpoint_steps=7)
        g = h.adjoint_backward(np.stack([p[2] for p in probs], axis=1))
    off = 0
    bath = {k: 0 for k in ARRAYS}
    for i, (inst, p) in enumerate(zip(insts, probs)):
        Xr, gr = inst.adjoint_gradient(p[0], 1e-15, nsteps, obs, p[1], p[2], checkpoint_steps=7)
        assert max_abs_rel(X[:, i], Xr) < 1e-12
        assert max_abs_rel(g["hamiltonian"][i], gr["hamiltonian"]) < 1e-10
        M = inst.matrices
        assert max_abs_rel(g["pf_same_sum"][off:off + M], gr["pf_same_sum"]) < 1e-10
        n = inst.state_size
        assert max_abs_rel(g["state"][off * n // M:(off + M) * n // M], gr["state"]) < 1e-10
        off += M
        for k in ("pf_same", "pf_circ", "cbk"):
            bath[k] = bath[k] + gr[k]
    for k in ("pf_same", "pf_circ", "cbk"):
        assert max_abs_rel(g[k], bath[k]) < 1e-10, k


def test_adjoint_rejects_fp32_and_masks(oracle_mod):
    _, inst = instance("excitonic7_d4", depth=2)
    with heom_amd.Handle.from_instance(inst, precision=heom_amd.HEOM_PRECISION_FP32) as h:
        with pytest.raises(heom_amd.HeomError, match="FP64"):
            h.rhs_adjoint()
    with heom_amd.Handle.from_instance(inst) as h:
        with pytest.raises(heom_amd.HeomError, match="record"):
            h.adjoint_backward(np.ones(3))
        m = np.ones(inst.matrices, np.uint8)
        m[-1] = 0
        h.set_mask(m)
        with pytest.raises(heom_amd.HeomError, match="unfiltered"):
            h.adjoint_forward(1e-15, 10, 1, np.eye(inst.states))


# ---- the adjoint-gradient workload --------------------------------------------------

SITE = 3


def params_vector(p):
    """8 site energies, 28 couplings (i < j), 8 lambda, 8 gamma."""
    H = np.real(p["hamiltonian"])
    iu = np.triu_indices(H.shape[0], 1)
    return np.concatenate([np.diag(H), H[iu], p["lambda"], p["gamma"]])


def grads_vector(pg):
    dH = np.real(pg["hamiltonian"])
    iu = np.triu_indices(dH.shape[0], 1)
    return np.concatenate([np.diag(dH), dH[iu] + dH.T[iu], pg["lambda"], pg["gamma"]])


def set_from_vector(host, v, S):
    H = np.diag(v[:S])
    iu = np.triu_indices(S, 1)
    H[iu] = v[S:S + len(iu[0])]
    H.T[iu] = v[S:S + len(iu[0])]
    B = host.baths
    host.set_parameters(H, v[-2 * B:-B], v[-B:])


@pytest.fixture(scope="module")
def workload():
    host = heom_amd.HostInstance(config_path("excitonic8_d6"))
    info = host.info
    N = host.states
    op = np.zeros((N, N))
    op[SITE, SITE] = 1.0
    nsteps, dt = int(info.steps), float(info.step_size)
    w = np.full(nsteps + 1, dt * 1e15)
    w[0] = w[-1] = 0.5 * dt * 1e15
    p0 = host.get_parameters()
    yield host, op, w, nsteps, dt, p0
    host.close()


def gpu_objective(host, op, w, nsteps, dt):
    with host.create_handle() as h:
        h.use_graphs(True)
        _, tr = h.propagate_observe(dt, nsteps, 1, op)
    return float(np.sum(w * np.real(tr[:, 0])))


def adjoint_gradient(host, op, w, nsteps, dt):
    with host.create_handle() as h:
        h.use_graphs(True)
        X = h.adjoint_forward(dt, nsteps, 1, op, checkpoint_steps=50)[:, 0]
        g = h.adjoint_backward(w, state=False)
    return float(np.sum(w * np.real(X))), g, host.parameter_gradient(g)


def test_workload_gradient_matches_finite_differences(oracle_mod, workload):
    host, op, w, nsteps, dt, p0 = workload
    host.set_parameters(p0["hamiltonian"], p0["lambda"], p0["gamma"])
    J, g, pg = adjoint_gradient(host, op, w, nsteps, dt)
    assert abs(J - gpu_objective(host, op, w, nsteps, dt)) <= 1e-13 * abs(J)
    grad = grads_vector(pg)
    v0 = params_vector(p0)
    S = host.sites
    fd = np.zeros_like(v0)
    hstep = 1e-3  # cm^-1
    for k in range(v0.size):
        vals = []
        for sgn in (1.0, -1.0):
            v = v0.copy()
            v[k] += sgn * hstep
            set_from_vector(host, v, S)
            vals.append(gpu_objective(host, op, w, nsteps, dt))
        fd[k] = (vals[0] - vals[1]) / (2 * hstep)
    set_from_vector(host, v0, S)
    scale = np.max(np.abs(grad))
    rel = np.abs(grad - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    names = ([f"eps{i}" for i in range(S)] + [f"J{i}{j}" for i, j in zip(*np.triu_indices(S, 1))]
             + [f"lambda{b}" for b in range(host.baths)] + [f"gamma{b}" for b in range(host.baths)])
    worst = int(np.argmax(rel))
    print(f"J = {J:.12e} fs; adjoint vs central differences (h = 1e-3 cm^-1): max rel {rel.max():.2e} "
          f"at {names[worst]} ({grad[worst]:.6e} vs {fd[worst]:.6e}); norm-wise "
          f"{np.linalg.norm(grad - fd) / np.linalg.norm(fd):.2e}", flush=True)
    assert rel.max() < 1e-6


def test_workload_gradient_matches_oracle_gradient(oracle_mod, workload):
    host, op, w, nsteps, dt, p0 = workload
    host.set_parameters(p0["hamiltonian"], p0["lambda"], p0["gamma"])
    J, g, pg = adjoint_gradient(host, op, w, nsteps, dt)
    cfg = O.parse_config(config_path("excitonic8_d6"))
    inst = O.instance_from_config(cfg)
    y0 = inst.state_with_top(cfg.rho_init)
    Xr, gr = inst.adjoint_gradient(y0, dt, nsteps, 1, op, w, checkpoint_steps=50)
    Jr = float(np.sum(w * np.real(Xr)))
    assert abs(J - Jr) <= 1e-10 * abs(Jr)
    bath = inst.bath_parameter_gradient(gr)
    ref = grads_vector({"hamiltonian": gr["hamiltonian"] * O.INVCM_TO_JOULE, "lambda": bath[0], "gamma": bath[1]})
    got = grads_vector(pg)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)))
    print(f"adjoint vs the oracle's FP64 gradient: max rel {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-6
    # the coefficient-array gradients themselves
    assert max_abs_rel(g["hamiltonian"][0], gr["hamiltonian"]) < 1e-10
    for k in ARRAYS:
        assert max_abs_rel(g[k], gr[k]) < 1e-10, k


def test_torch_wrapper_gradients(oracle_mod, workload):
    host, op, w, nsteps, dt, p0 = workload
    host.set_parameters(p0["hamiltonian"], p0["lambda"], p0["gamma"])
    _, _, pg = adjoint_gradient(host, op, w, nsteps, dt)
    system = System(config_path("excitonic8_d6"))
    try:
        P = system.parameters()
        eps = torch.diagonal(P["hamiltonian"]).detach().clone().requires_grad_(True)
        S = system.sites
        iu = torch.triu_indices(S, S, 1)
        cpl = P["hamiltonian"][iu[0], iu[1]].detach().clone().requires_grad_(True)
        H = torch.diag(eps)
        H = H.index_put((iu[0], iu[1]), cpl).index_put((iu[1], iu[0]), cpl)
        J = system.population_integral(SITE, H, P["lam"], P["gamma"])
        J.backward()
        got = np.concatenate([eps.grad.numpy(), cpl.grad.numpy(), P["lam"].grad.numpy(), P["gamma"].grad.numpy()])
        want = grads_vector(pg)
        assert max_abs_rel(got, want) < 1e-12
        # a second forward with other parameters before the first backward
        J1 = system.population_integral(SITE, H.detach(), P["lam"].detach() * 1.1, P["gamma"].detach())
        lam = P["lam"].detach().clone().requires_grad_(True)
        J2 = system.population_integral(SITE, H.detach(), lam, P["gamma"].detach())
        J2.backward()
        assert max_abs_rel(lam.grad.numpy(), pg["lambda"]) < 1e-12
        assert float(J1.detach()) != float(J2.detach())
    finally:
        system.close()
