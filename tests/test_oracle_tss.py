"""The oracle's thermal state search (oracle_rhs_tss, oracle_tss_step, thermal_state_setup,
thermal_state_search) pinned on the CPU against what the mathematics fixes:

* the start is the Boltzmann state exp(-beta H) / Z (scipy's expm, not an eigenbasis);
* as the system-bath coupling vanishes the Boltzmann state becomes a fixed point;
* the converged hierarchy is the null vector of the HEOM generator, written as a dense
  matrix by brute-force ADO enumeration (test_oracle.dense_liouvillian) and solved with
  numpy's SVD, with the top's trace fixed to one -- a dropped term, a wrong sign or a
  transposed C in the search would converge elsewhere or not at all;
* the f_TSS evaluation is eps y plus the ordinary RHS without its commutator and S_u terms.
The reference's Eigen ComplexEigenSolver is an un-vendored dependency; the iteration does
not depend on the order or signs of the eigenvectors of a real symmetric H (both cancel
in C^T [(C k C^T) / D] C), so numpy's eigh stands in for it."""
import numpy as np
import pytest
import scipy.linalg

import oracle as O
from helpers import FMO_DIR
from test_oracle import dense_liouvillian, small_instance, to_dense

TSS_CFG = FMO_DIR + "/fmo_thermal_state_77K.cfg"


def test_config_task_and_checks(tmp_path):
    cfg = O.parse_config(TSS_CFG)
    assert cfg.task == "thermal_state_search"
    assert cfg.tss_delta == 1e-8 and cfg.tss_max_steps == 5000 and cfg.step_size == 1.0
    bad = tmp_path / "bad.cfg"
    bad.write_text(open(TSS_CFG).read().replace("step_size=1.0", "step_size=2.0"))
    with pytest.raises(O.ConfigError, match="must equal 1.0"):
        O.parse_config(str(bad))
    with pytest.raises(O.ConfigError, match="wrong program_task"):
        O.parse_config(TSS_CFG, "population_dynamics")


@pytest.mark.parametrize("T", [77.0, 300.0])
def test_setup_is_the_boltzmann_state(T):
    inst, H = small_instance(5, 2, 2, [[0], [1], [2], [3], [4]], seed=3, T=T)
    s = O.thermal_state_setup(H, T)
    beta = 1.0 / (O.KB * T)
    Hs = H.real - np.min(np.linalg.eigvalsh(H.real)) * np.eye(5)  # shifted: the ratio is unchanged
    rho = scipy.linalg.expm(-beta * Hs)
    rho /= np.trace(rho)
    assert np.allclose(s["top"], rho, rtol=0, atol=1e-13)
    C = s["c"]
    assert np.allclose(C @ C.T, np.eye(5), atol=1e-14)
    assert np.allclose(C @ H @ C.T, np.diag(s["eigenvalues"]), rtol=0, atol=1e-13 * np.abs(H).max())
    assert np.array_equal(s["pseudo_hamiltonian"], np.full((5, 5), 1e14))
    with pytest.raises(ValueError, match="real symmetric"):
        O.thermal_state_setup(H + 1j * 1e-22 * (np.eye(5, k=1) - np.eye(5, k=-1)), T)


def test_ground_state_is_kept_empty():
    _, H = small_instance(3, 1, 1, [[0], [1], [2]], seed=1)
    Hg = O.state_hamiltonian(H, "with_ground_state")
    s = O.thermal_state_setup(Hg, 300.0, "with_ground_state")
    assert abs(s["top"][0, 0]) < 1e-300 and abs(np.trace(s["top"]) - 1) < 1e-14
    assert np.isclose(np.max(s["eigenvalues"].real), O.TSS_GROUND_STATE_ENERGY)


def tss_instance(inst, H, T):
    s = O.thermal_state_setup(H, T)
    inst.set_hamiltonian(s["pseudo_hamiltonian"])
    return s


def test_rhs_tss_is_eps_y_plus_the_bath_terms():
    inst, H = small_instance(3, 2, 2, [[0], [1], [2]], seed=4)
    rng = np.random.default_rng(0)
    y = rng.normal(size=inst.state_size)
    s = tss_instance(inst, H, 300.0)
    got = inst.rhs_tss(y)
    # the ordinary RHS with H = 0 and S_u = 0, plus eps y
    inst.set_hamiltonian(np.zeros_like(H))
    pss = inst.pf_same_sum.copy()
    inst.pf_same_sum[:] = 0.0
    want = inst.rhs(y) + s["epsilon"].real * y
    inst.pf_same_sum[:] = pss
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_weakly_coupled_boltzmann_state_is_a_fixed_point():
    # lambda -> 0 (lambda = 0 itself is the reference's 0 * inf in the minus term): the
    # top moves by O(lambda), here a relative 1e-12
    inst, H = small_instance(4, 2, 2, [[0], [1], [2], [3]], seed=5, lam_scale=1e-12)
    s = tss_instance(inst, H, 77.0)
    y = inst.state_with_top(s["top"])
    inst.tss_step(y, s, 3)
    assert np.max(np.abs(inst.top(y) - s["top"])) < 1e-10


@pytest.mark.parametrize("seed,T", [(6, 300.0), (7, 77.0)])
def test_converged_search_is_the_null_vector_of_the_generator(seed, T):
    inst, H = small_instance(3, 1, 2, [[0], [1], [2]], seed=seed, T=T)
    L, ados, _ = dense_liouvillian(inst, H)
    # the stationary hierarchy: the right singular vector of the smallest singular value,
    # scaled to a top of trace one
    _, sv, vh = np.linalg.svd(L)
    assert sv[-1] < 1e-10 * sv[0] and sv[-2] > 1e-6 * sv[0]  # a one-dimensional null space
    x = vh[-1].conj()
    n2 = inst.states ** 2
    x = x / np.trace(x[:n2].reshape(inst.states, inst.states))
    s = tss_instance(inst, H, T)

    def search(setup):
        y = inst.state_with_top(setup["top"])
        for _ in range(200):
            prev = y.copy()
            inst.tss_step(y, setup, 100)
            if np.max(np.abs(y - prev)) < 1e-15:
                break
        return y

    y = search(s)
    got = to_dense(inst, y, ados)
    err = np.max(np.abs(got - x)) / np.max(np.abs(x))
    print(f"search vs null vector: {err:.1e}")
    assert err < 1e-9
    # a planted sign error in the denominator, i (E_j - E_i) / hbar, ends elsewhere
    bad = to_dense(inst, search(dict(s, eigenvalues=-s["eigenvalues"])), ados)
    assert np.max(np.abs(bad - x)) > 1e-4 * np.max(np.abs(x))
    # the populations are not the Boltzmann ones: the bath changes the equilibrium
    assert np.max(np.abs(np.diag(got[:n2].reshape(3, 3)) - np.diag(s["top"]))) > 1e-6


def test_fmo_search_converges_like_the_reference_loop():
    cfg = O.parse_config(TSS_CFG)
    tops, y, it = O.thermal_state_search(cfg)
    assert cfg.steps < it < cfg.tss_max_steps
    n1 = np.abs(tops).sum(axis=(1, 2))
    assert abs(n1[-1] - n1[-2]) < cfg.tss_delta and abs(n1[-2] - n1[-3]) >= cfg.tss_delta
    assert abs(np.trace(tops[-1]) - 1) < 1e-10
    assert np.allclose(tops[-1], tops[-1].conj().T, atol=1e-12)
    # stationary for the true generator, relative to the Boltzmann start
    inst = O.instance_from_config(cfg)
    r0 = np.abs(inst.rhs(inst.state_with_top(tops[0]))).max()
    assert np.abs(inst.rhs(y)).max() < 1e-6 * r0
