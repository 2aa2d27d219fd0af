"""The product's applications end to end on the GPU against the oracle.

heom_population_dynamics, heom_linear_absorption and heom_thermal_state_search
(dm-heom_amd/bin, the counterparts of dm-heom/src/app_population_dynamics.cpp,
app_linear_absorption.cpp and app_thermal_state_search.cpp) read a reference config
file, build the instance with the C++ host library, run the HIP kernels and write
observation files.  Their numbers are compared with the oracle's run of the same config:
populations and absorption traces within 1e-10 relative to the largest value.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import FMO_DIR, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "dm-heom_amd", "bin")


def write_cfg(src, dst, **replace):
    text = open(src).read()
    for key, value in replace.items():
        text, n = re.subn(rf"^{key}=.*$", f"{key}={value}", text, count=1, flags=re.M)
        assert n == 1, key
    with open(dst, "w") as f:
        f.write(text)
    return dst


def run_app(app, cfg_path, cwd, *extra):
    exe = os.path.join(BIN, app)
    assert os.path.exists(exe), "build the applications first (make -C dm-heom_amd)"
    r = subprocess.run([exe, "--no-progress", *extra, cfg_path, cfg_path], cwd=cwd, capture_output=True,
                       text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r


def read_table(path):
    with open(path) as f:
        header = f.readline().split()
        rows = [list(map(float, line.split())) for line in f if line.strip()]
    return header, np.array(rows)


def test_population_dynamics_app_fmo(oracle_mod, tmp_path):
    """FMO example (7 sites, 14 modes, depth 3), 2000 steps observed every 5."""
    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_population_dynamics_77K.cfg"), tmp_path / "pd.cfg",
                         steps="2000", observations="{(matrix_diagonal, pd_diag.dat), (matrix_trace_id, pd_trace.dat)}")
    run_app("heom_population_dynamics", str(cfg_path), tmp_path)
    header, got = read_table(tmp_path / "pd_diag.dat")
    assert header == ["time"] + [f"site_{i}" for i in range(7)]

    cfg = O.parse_config(str(cfg_path))
    inst = O.instance_from_config(cfg)
    y = inst.state_with_top(cfg.rho_init)
    tops = inst.propagate(y, cfg.step_size, cfg.steps, cfg.observe_steps)
    pops = np.real(np.diagonal(tops, axis1=1, axis2=2))
    times = np.arange(tops.shape[0]) * cfg.observe_steps * cfg.step_size
    assert got.shape == (tops.shape[0], 8)
    assert np.allclose(got[:, 0], times, rtol=1e-15, atol=0)
    err = np.max(np.abs(got[:, 1:] - pops)) / np.max(np.abs(pops))
    print(f"FMO populations: max rel deviation {err:.3e}", flush=True)
    assert err < 1e-10
    _, tr = read_table(tmp_path / "pd_trace.dat")
    assert np.max(np.abs(tr[:, 1] - 1.0)) < 1e-12


def test_hierarchy_complete_observation(oracle_mod, tmp_path):
    """hierarchy_complete (dm-heom/include/heom/hierarchy_complete_observer.hpp:42-82): every
    ADO in uid order at every observation, FMO for 20 steps observed every 10, against the
    oracle's state (1e-10 relative to the largest value), and the reference's column names."""
    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_population_dynamics_77K.cfg"), tmp_path / "pd.cfg",
                         steps="20", observe_steps="10",
                         observations="{(hierarchy_complete, pd_all.dat), (matrix_diagonal, pd_diag.dat)}")
    run_app("heom_population_dynamics", str(cfg_path), tmp_path)
    header, got = read_table(tmp_path / "pd_all.dat")
    cfg = O.parse_config(str(cfg_path))
    inst = O.instance_from_config(cfg)
    N, M = inst.states, inst.matrices
    assert header[:4] == ["time", "mat_0_elem_0_0_real", "mat_0_elem_0_0_imag", "mat_0_elem_0_1_real"]
    assert header[-1] == f"mat_{M - 1}_elem_{N - 1}_{N - 1}_imag" and len(header) == 1 + M * N * N * 2
    y = inst.state_with_top(cfg.rho_init)
    rows = [y.copy()]
    work = np.empty(5 * inst.state_size)
    for n in range(1, cfg.steps + 1):
        inst.rk4_step(y, cfg.step_size, work)
        if n % cfg.observe_steps == 0:
            rows.append(y.copy())
    want = np.array(rows)
    assert got.shape == (3, 1 + M * N * N * 2)
    assert np.allclose(got[:, 0], np.arange(3) * 10 * cfg.step_size, rtol=1e-15, atol=0)
    err = np.max(np.abs(got[:, 1:] - want)) / np.max(np.abs(want))
    print(f"hierarchy_complete: {M} ADOs x 3 observations, max rel deviation {err:.3e}", flush=True)
    assert err < 1e-10
    _, diag = read_table(tmp_path / "pd_diag.dat")
    assert diag.shape == (3, 1 + N)


@pytest.mark.parametrize("stepper", ["rk_heun", "rk_kutta3", "rk_rk4_38"])
def test_population_dynamics_app_steppers(oracle_mod, tmp_path, stepper):
    """solver.stepper_type other than rk_rk4: the app steps through the per-evaluation
    interface; FMO, 300 steps observed every 5, against the same tableau on the host with
    the oracle's RHS (1e-10)."""
    from heom_amd import steppers

    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_population_dynamics_77K.cfg"), tmp_path / "pd.cfg",
                         steps="300", stepper_type=stepper, observations="{(matrix_diagonal, pd_diag.dat)}")
    run_app("heom_population_dynamics", str(cfg_path), tmp_path)
    _, got = read_table(tmp_path / "pd_diag.dat")
    cfg = O.parse_config(str(cfg_path))
    inst = O.instance_from_config(cfg)
    y = inst.state_with_top(cfg.rho_init)
    pops = []
    for n in range(cfg.steps + 1):
        if n % cfg.observe_steps == 0:
            top = y[:inst.states * inst.states * 2].reshape(inst.states, inst.states, 2)
            pops.append(np.diagonal(top[..., 0]).copy())
        if n < cfg.steps:
            y = steppers.numpy_step(inst.rhs, y, n * cfg.step_size, cfg.step_size, stepper)
    pops = np.array(pops)
    assert got.shape == (pops.shape[0], 8)
    err = np.max(np.abs(got[:, 1:] - pops)) / np.max(np.abs(pops))
    print(f"{stepper}: FMO populations, max rel deviation {err:.3e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("rotating", [False, True])
def test_linear_absorption_app_fmo(oracle_mod, tmp_path, rotating):
    """FMO linear absorption (8 states with the ground state, 3 tensor components),
    600 steps; with and without the rotating frame (E0 = 1400 cm^-1)."""
    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_linear_absorption_77K.cfg"), tmp_path / "la.cfg",
                         steps="600", observe_steps="2",
                         observations="{(matrix_trace_static_fluorescence, la.dat)}")
    extra = ["--rotating-frame", "1400"] if rotating else []
    run_app("heom_linear_absorption", str(cfg_path), tmp_path, *extra)
    header, got = read_table(tmp_path / "la.dat")
    assert header == ["time", "trace_real", "trace_imag"]

    cfg = O.parse_config(str(cfg_path))
    e0 = 1400 * O.INVCM_TO_JOULE if rotating else 0.0
    times, ref = O.linear_absorption(cfg, rotating_frame=e0)
    assert got.shape == (len(times), 3)
    assert np.allclose(got[:, 0], times, rtol=1e-15, atol=0)
    val = got[:, 1] + 1j * got[:, 2]
    err = np.max(np.abs(val - ref)) / np.max(np.abs(ref))
    print(f"FMO absorption trace (rotating={rotating}): max rel deviation {err:.3e}", flush=True)
    assert err < 1e-10


def test_thermal_state_search_app_fmo(oracle_mod, tmp_path):
    """The FMO thermal-state example as shipped (500 steps at least, delta 1e-8): one row
    per iteration, the same iteration count and converged populations as the oracle's
    search loop."""
    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_thermal_state_77K.cfg"), tmp_path / "tss.cfg",
                         observations="{(matrix_diagonal, tss_diag.dat), (matrix_trace_id, tss_trace.dat)}")
    r = run_app("heom_thermal_state_search", str(cfg_path), tmp_path)
    assert "converged" in r.stdout and "not converged" not in r.stdout
    header, got = read_table(tmp_path / "tss_diag.dat")
    assert header == ["time"] + [f"site_{i}" for i in range(7)]
    cfg = O.parse_config(str(cfg_path))
    tops, _, it = O.thermal_state_search(cfg)
    pops = np.real(np.diagonal(tops, axis1=1, axis2=2))
    assert got.shape == (it + 1, 8)
    assert np.allclose(got[:, 0], np.arange(it + 1) * cfg.observe_steps, rtol=0, atol=0)
    err = np.max(np.abs(got[:, 1:] - pops))
    print(f"FMO thermal state: {it} iterations, max population deviation {err:.1e}", flush=True)
    assert err < 1e-10
    _, tr = read_table(tmp_path / "tss_trace.dat")
    assert np.max(np.abs(tr[:, 1] - 1.0)) < 1e-9


def test_thermal_state_search_app_filtered_with_norms(oracle_mod, tmp_path):
    """The app with the single-excitation filter (first layer 1) and a hierarchy_norm
    observation: the last row's populations and ADO norms against the oracle's search
    with the host library's mask, replayed for as many iterations as the app wrote."""
    heom_amd = pytest.importorskip("heom_amd")
    cfg_path = write_cfg(os.path.join(FMO_DIR, "fmo_thermal_state_77K.cfg"), tmp_path / "tssf.cfg",
                         observations="{(matrix_diagonal, f_diag.dat), (hierarchy_norm, f_norm.dat)}",
                         strategy="single_excitation", first_layer="1", max_steps="200", steps="100")
    r = run_app("heom_thermal_state_search", str(cfg_path), tmp_path)
    assert "thermal state search:" in r.stdout
    _, got = read_table(tmp_path / "f_diag.dat")
    cfg = O.parse_config(str(cfg_path))
    inst = O.instance_from_config(cfg)
    mask = heom_amd.HostInstance(str(cfg_path), task=heom_amd.TASK_THERMAL_STATE_SEARCH).mask()
    s = O.thermal_state_setup(inst.hamiltonian_matrix(), cfg.baths_temperature)
    inst.set_hamiltonian(s["pseudo_hamiltonian"])
    inst.mask[:] = mask
    y = inst.state_with_top(s["top"])
    rows = got.shape[0] - 1
    inst.tss_step(y, s, rows * cfg.observe_steps)
    pops = np.real(np.diag(inst.top(y)))
    err = np.max(np.abs(got[-1, 1:] - pops))
    print(f"filtered FMO thermal state: {rows} iterations, population deviation {err:.1e}", flush=True)
    assert err < 1e-10
    _, norms = read_table(tmp_path / "f_norm.dat")
    assert norms.shape == (got.shape[0], inst.matrices + 1)
    ref = inst.norms(y)
    nerr = np.max(np.abs(norms[-1, 1:] - ref)) / np.max(ref)
    assert nerr < 1e-10
