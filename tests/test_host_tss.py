"""Host side of the thermal state search (libheom_host.so, no GPU): the Jacobi eigen
decomposition against numpy's eigh (random, degenerate and badly scaled matrices), the
set-up (C, eigenvalues, Boltzmann top, epsilon) against the oracle's, and the config
task with its checks."""
import numpy as np
import pytest

import oracle as O
from helpers import FMO_DIR, config_path

heom_amd = pytest.importorskip("heom_amd")

TSS_CFG = FMO_DIR + "/fmo_thermal_state_77K.cfg"


def test_fmo_setup_matches_the_oracle():
    host = heom_amd.HostInstance(TSS_CFG, task=heom_amd.TASK_THERMAL_STATE_SEARCH)
    s = host.tss_setup()
    cfg = O.parse_config(TSS_CFG)
    inst = O.instance_from_config(cfg)
    H = inst.hamiltonian_matrix()
    r = O.thermal_state_setup(H, cfg.baths_temperature)
    assert np.max(np.abs(s["top"] - r["top"])) < 1e-14
    assert np.max(np.abs(host.top() - r["top"])) < 1e-14  # the instance starts there
    C = s["c"]
    assert np.max(np.abs(C @ C.T - np.eye(7))) < 1e-14
    assert np.max(np.abs(C @ H @ C.T - np.diag(s["eigenvalues"]))) < 1e-14 * np.abs(H).max()
    assert np.allclose(np.sort(s["eigenvalues"].real), np.sort(r["eigenvalues"].real), rtol=1e-14, atol=0)
    assert s["epsilon"] == 1e14
    assert host.info.tss_delta == 1e-8 and host.info.tss_max_steps == 5000


@pytest.mark.parametrize("n,kind", [(1, "random"), (2, "random"), (8, "random"), (24, "random"), (32, "random"),
                                    (6, "degenerate"), (5, "identity"), (7, "ground state")])
def test_jacobi_eigen_decomposition(n, kind):
    rng = np.random.default_rng(n)
    if kind == "identity":
        H = np.eye(n) * 2e-19
    elif kind == "degenerate":
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        H = Q @ np.diag([1.0, 1.0, 1.0, 2.0, 2.0, 3.0]) @ Q.T * 1e-20
    else:
        H = rng.normal(size=(n, n)) * 1e-21 + np.diag(rng.uniform(2.3e-19, 2.5e-19, n))
        if kind == "ground state":  # the reference's ground-state energy far above the rest
            H[0, :] = H[:, 0] = 0.0
            H[0, 0] = 1e5
    H = (H + H.T) / 2
    E, V = heom_amd.symmetric_eigen(H)
    w = np.linalg.eigvalsh(H)
    scale = np.abs(w).max()
    assert np.allclose(np.sort(E), w, rtol=0, atol=1e-14 * scale)
    assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-13
    off = V.T @ H @ V - np.diag(E)
    assert np.max(np.abs(off)) < 1e-13 * scale
    if n > 1:
        with pytest.raises(heom_amd.HeomError, match="not symmetric"):
            heom_amd.symmetric_eigen(H + np.triu(np.ones((n, n)), 1) * 1e-25)


def test_tss_task_checks(tmp_path):
    bad = tmp_path / "bad.cfg"
    bad.write_text(open(TSS_CFG).read().replace("step_size=1.0", "step_size=0.5"))
    with pytest.raises(heom_amd.ConfigError, match="must equal 1.0"):
        heom_amd.HostInstance(str(bad), task=heom_amd.TASK_THERMAL_STATE_SEARCH)
    with pytest.raises(heom_amd.ConfigError, match="wrong program_task"):
        heom_amd.HostInstance(TSS_CFG, task=heom_amd.TASK_POPULATION_DYNAMICS)
    cplx = tmp_path / "c.cfg"
    cplx.write_text(open(TSS_CFG).read().replace("{1410.000, -87.70000,", "{1410.000, (-87.70000, 1.0),")
                    .replace("{-87.70000, 1530.000,", "{(-87.70000, -1.0), 1530.000,"))
    with pytest.raises(heom_amd.ConfigError, match="real symmetric"):
        heom_amd.HostInstance(str(cplx), task=heom_amd.TASK_THERMAL_STATE_SEARCH)
