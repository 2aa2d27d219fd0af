"""Host side of the adjoint (libheom_host.so, no GPU): parameter overrides recompute the
prefactors exactly as the oracle's precompute does for the same parameters, and the
analytic chain rule of heom_host_parameter_gradient (coefficient-array gradients ->
dJ/d lambda, gamma, Omega, H) agrees with central differences of the oracle's
precompute."""
import numpy as np
import pytest

import oracle as O
from helpers import FMO_DIR, config_path

heom_amd = pytest.importorskip("heom_amd")

KEYS = ("pf_same_sum", "pf_same", "pf_circ", "cbk")


def oracle_instance(path, task, params, depth=0):
    cfg = O.parse_config(path)
    if depth:
        cfg.ado_depth = depth
    cfg.baths_lambda = params["lambda"] * O.INVCM_TO_JOULE
    cfg.baths_invnu = O.HBAR / (params["gamma"] * O.INVCM_TO_JOULE)
    cfg.baths_omega = params["omega"] * O.INVCM_TO_JOULE / O.HBAR
    cfg.hamiltonian = params["hamiltonian"] * O.INVCM_TO_JOULE
    mode = "identity" if task == heom_amd.TASK_POPULATION_DYNAMICS else "with_ground_state"
    return O.instance_from_config(cfg, mode)


CASES = [(config_path("excitonic8_d6"), heom_amd.TASK_POPULATION_DYNAMICS, 3),
         (config_path("excitonic8_d8"), heom_amd.TASK_POPULATION_DYNAMICS, 3),
         (FMO_DIR + "/fmo_linear_absorption_77K.cfg", heom_amd.TASK_LINEAR_ABSORPTION, 0)]


@pytest.mark.parametrize("path,task,depth", CASES)
def test_set_parameters_recompute_the_prefactors(oracle_mod, path, task, depth):
    host = heom_amd.HostInstance(path, task=task, depth=depth)
    p = host.get_parameters()
    rng = np.random.default_rng(1)
    new = {
        "hamiltonian": p["hamiltonian"] + np.diag(rng.normal(0, 20, host.sites)),
        "lambda": p["lambda"] * rng.uniform(0.7, 1.3, host.baths),
        "gamma": p["gamma"] * rng.uniform(0.7, 1.3, host.baths),
        "omega": rng.uniform(-50, 50, host.baths),
    }
    host.set_parameters(new["hamiltonian"], new["lambda"], new["gamma"], new["omega"])
    got = host.get_parameters()
    for k in new:
        assert np.allclose(got[k], new[k], rtol=1e-14, atol=1e-12), k
    # the oracle gets the parameters back in cm^-1 and converts them again (one rounding
    # apart from the host's SI values), so the comparison is to rounding, not bitwise
    ref = oracle_instance(path, task, got, depth)
    arrays = host.arrays()
    for k in KEYS + ("hamiltonian",):
        want = np.asarray(getattr(ref, k)).reshape(-1)
        assert np.max(np.abs(arrays[k] - want)) <= 1e-13 * np.max(np.abs(want)), k


@pytest.mark.parametrize("path,task,depth", CASES)
def test_parameter_gradient_chain_matches_differences_of_precompute(oracle_mod, path, task, depth):
    host = heom_amd.HostInstance(path, task=task, depth=depth)
    p = host.get_parameters()
    rng = np.random.default_rng(2)
    p["omega"] = rng.uniform(-40, 40, host.baths)  # exercise Im nu
    host.set_parameters(omega=p["omega"])
    ref = oracle_instance(path, task, p, depth)
    N = host.states
    grads = {k: rng.normal(size=np.asarray(getattr(ref, k)).size // 2)
             + 1j * rng.normal(size=np.asarray(getattr(ref, k)).size // 2) for k in KEYS}
    grads["hamiltonian"] = (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))[None]
    got = host.parameter_gradient(grads)
    want = ref.bath_parameter_gradient(grads)
    for row, key in enumerate(("lambda", "gamma", "omega")):
        err = np.max(np.abs(got[key] - want[row])) / np.max(np.abs(want[row]))
        print(f"{key}: max rel deviation {err:.2e}")
        assert err < 1e-7, key
    off = N - host.sites
    assert np.allclose(got["hamiltonian"], grads["hamiltonian"][0][off:, off:] * O.INVCM_TO_JOULE, rtol=1e-15)


def test_set_parameters_rejects_nonpositive_gamma():
    host = heom_amd.HostInstance(config_path("excitonic8_d6"), depth=2)
    with pytest.raises(heom_amd.HeomError, match="gamma"):
        host.set_parameters(gamma=np.zeros(host.baths))
