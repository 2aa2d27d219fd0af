"""Pins the oracle's bath prefactors and its linear-absorption pipeline by physics.

The bath of every state is a Drude-Lorentz bath, J(w) = 2 lambda gamma w / (w^2 + gamma^2).
Its correlation function, in the units the hierarchy uses (C(t)/hbar, J/s), is

    c(t) = (1/pi) int_0^inf J(w) [coth(beta hbar w / 2) cos(w t) - i sin(w t)] dw.

The oracle's prefactors (dm-heom/src/heom/instance.cpp:155-262) claim that

* sum_k c_k exp(-nu_k t) approximates Re c(t), better with every Pade term K;
* hbar Phi_b exp(-gamma t) = -Im c(t) exactly;
* hbar Delta_b is the zero-frequency weight the kept exponentials miss:
  int_0^inf [Re c(t) - sum_k c_k exp(-nu_k t)] dt = S(0)/2 - sum_k c_k/nu_k, where
  S(w) = J(w) coth(beta hbar w / 2), because int_0^inf cos(w t) dt puts half of a delta
  function at w = 0.

Here c(t) comes from adaptive Fourier quadrature (QUADPACK QAWF through scipy) of J(w)
itself, cross-checked against the textbook Matsubara series, so no prefactor formula is
reused.  The linear-absorption pipeline (dipole matrices D-/D+, the trace Tr(D- sigma_0),
the tensor average and the rotating frame; dm-heom/src/app_linear_absorption.cpp:88-163,
dm-heom/include/heom/dipole_matrix.hpp:30-74) is pinned three ways:

* lambda -> 0: the average of mu_c^T exp(-i H t / hbar) mu_c over the tensor components,
  with mu_c built here from the config's directions and strengths and exp by scipy;
* a two-level monomer (one site and the ground state) with a Drude bath: the hierarchy
  reproduces the exact second-cumulant line shape exp(-i eps t / hbar - g(t)), with
  g(t) = sum_k (c_k/hbar) F(nu_k, t) + Delta t - i Phi F(nu_0, t),
  F(v, t) = (exp(-v t) + v t - 1) / v^2, as the depth grows;
* the rotating frame is an exact transformation, so with a bath the plain and rotating
  runs differ only by RK4 truncation error, which must fall as dt^4.

Each check also runs on a planted mistake (a factor 2 in c_k, a sign flipped in D+ or in
the frame phase) and must then fail.
"""
import os

import numpy as np
import pytest
import scipy.linalg
from scipy import integrate

import oracle as O
from helpers import FMO_DIR

CM = O.INVCM_TO_JOULE


# ---------------------------------------------------------------------------------
# the Drude-Lorentz correlation function by quadrature
# ---------------------------------------------------------------------------------

def re_c_quadrature(lam, gamma, T, t):
    """(1/pi) int J(w) coth(beta hbar w/2) cos(w t) dw, in x = w/gamma:
    (2 lambda gamma/pi) int x coth(a x)/(1+x^2) cos(gamma t x) dx, a = beta hbar gamma/2.
    coth = 1 + 2/(e^{2ax}-1): the first part decays as 1/x (QAWF), the second
    exponentially (plain adaptive quadrature on a finite range)."""
    a = O.HBAR * gamma / (2 * O.KB * T)
    tau = gamma * t
    slow, _ = integrate.quad(lambda x: x / (1 + x * x), 0, np.inf, weight="cos", wvar=tau, limlst=200)

    def fast(x):
        return 2 * x / ((1 + x * x) * np.expm1(2 * a * x)) * np.cos(tau * x) if x > 0 else 1.0 / a

    quick, _ = integrate.quad(fast, 0, 60 / a + 50, limit=400, epsabs=0, epsrel=1e-13)
    return 2 * lam * gamma / np.pi * (slow + quick)


def im_c_quadrature(lam, gamma, t):
    """-(1/pi) int J(w) sin(w t) dw = -(2 lambda gamma/pi) int x sin(gamma t x)/(1+x^2) dx."""
    val, _ = integrate.quad(lambda x: x / (1 + x * x), 0, np.inf, weight="sin", wvar=gamma * t, limlst=200)
    return -2 * lam * gamma / np.pi * val


def re_c_matsubara(lam, gamma, T, t, kmax=200000):
    """The textbook Matsubara series of the same function (Ishizaki and Tanimura 2005):
    lambda gamma cot(beta hbar gamma/2) e^{-gamma t}
      + (4 lambda gamma/(beta hbar)) sum_k nu_k e^{-nu_k t}/(nu_k^2 - gamma^2), nu_k = 2 pi k/(beta hbar)."""
    bh = O.HBAR / (O.KB * T)
    nu = 2 * np.pi * np.arange(1, kmax + 1) / bh
    return (lam * gamma / np.tan(bh * gamma / 2) * np.exp(-gamma * t)
            + 4 * lam * gamma / bh * np.sum(nu * np.exp(-nu * t) / (nu ** 2 - gamma ** 2)))


def spectral_zero_weight(lam, gamma, T):
    """S(0)/2 = lim_{w->0} J(w) coth(beta hbar w/2) / 2, evaluated at w = 1e-7 gamma."""
    w = 1e-7 * gamma
    return 0.5 * 2 * lam * gamma * w / (w * w + gamma * gamma) / np.tanh(O.HBAR * w / (2 * O.KB * T))


def oracle_bath(lam, invnu, T, K):
    """The oracle's c_k, nu_k, Delta, Phi for one bath with K terms (nu_k read back from
    S_n of the unit tuples, S_{e_m} = -nu_m)."""
    g = O.build_graph(K, 1)
    inst = O.Instance(states=1, matsubaras=K, baths=1, coupling=np.array([[0]], dtype=np.int32),
                      hamiltonian=np.zeros((1, 1), complex), graph=g, lambda_=np.array([lam]),
                      invnu=np.array([invnu]), omega=np.zeros(1), temperature=T)
    c = inst.cbk[0::2] + 1j * inst.cbk[1::2]
    S = inst.pf_same_sum[0::2] + 1j * inst.pf_same_sum[1::2]
    index = {tuple(t): u for u, t in enumerate(g.tuples)}
    nu = np.array([-S[index[tuple(1 if k == m else 0 for k in range(K))]] for m in range(K)])
    return {"c": c, "nu": nu, "delta": complex(inst.pf_same[0], inst.pf_same[1]),
            "phi": complex(inst.pf_circ[0], inst.pf_circ[1])}


def decomposition_error(bath, lam, gamma, T, ts):
    ref = np.array([re_c_quadrature(lam, gamma, T, t) for t in ts])
    approx = np.array([np.sum(bath["c"] * np.exp(-bath["nu"] * t)).real for t in ts])
    return float(np.max(np.abs(approx - ref)) / np.max(np.abs(ref)))


def delta_error(bath, lam, gamma, T):
    want = (spectral_zero_weight(lam, gamma, T) - np.sum(bath["c"] / bath["nu"]).real) / O.HBAR
    return abs(bath["delta"] - want) / abs(want)


def phi_error(bath, lam, gamma, ts):
    ref = np.array([im_c_quadrature(lam, gamma, t) for t in ts])
    got = -O.HBAR * bath["phi"] * np.exp(-bath["nu"][0] * ts)
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


BATHS = [(35.0, 50.0, 300.0), (35.0, 50.0, 77.0), (100.0, 20.0, 150.0)]  # lambda cm^-1, 1/gamma fs, T K


@pytest.mark.parametrize("lam_cm,invnu_fs,T", BATHS)
def test_quadrature_matches_matsubara_series(lam_cm, invnu_fs, T):
    """Two independent evaluations of Re c(t) agree: the quadrature is trustworthy."""
    lam, gamma = lam_cm * CM, 1 / (invnu_fs * 1e-15)
    ts = (5e-15, 20e-15, 100e-15, 300e-15)
    q = np.array([re_c_quadrature(lam, gamma, T, t) for t in ts])
    m = np.array([re_c_matsubara(lam, gamma, T, t) for t in ts])
    assert np.max(np.abs(q - m)) <= 1e-9 * np.max(np.abs(m)), (q, m)


@pytest.mark.parametrize("lam_cm,invnu_fs,T", BATHS)
def test_exponentials_converge_to_the_correlation_function(oracle_mod, lam_cm, invnu_fs, T):
    """sum_k c_k e^{-nu_k t} -> Re c(t) on t in [10, 300] fs as K grows: the error falls
    with every added Pade term and reaches the size of the neglected fastest pole."""
    lam, gamma = lam_cm * CM, 1 / (invnu_fs * 1e-15)
    ts = np.linspace(10e-15, 300e-15, 30)
    errs = [decomposition_error(oracle_bath(lam, invnu_fs * 1e-15, T, K), lam, gamma, T, ts) for K in range(1, 6)]
    assert all(b < a / 3 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < {300.0: 1e-8, 77.0: 1e-3, 150.0: 1e-4}[T], errs


def test_exponentials_fail_with_a_planted_factor(oracle_mod):
    lam, gamma, T = 35 * CM, 1 / 50e-15, 300.0
    ts = np.linspace(10e-15, 300e-15, 30)
    good = oracle_bath(lam, 50e-15, T, 4)
    base = decomposition_error(good, lam, gamma, T, ts)
    assert base < 1e-5
    for k in range(4):  # the fastest kept term has decayed to e^-8 at 10 fs, the slowest not at all
        bad = dict(good, c=good["c"].copy())
        bad["c"][k] *= 2
        assert decomposition_error(bad, lam, gamma, T, ts) > 30 * base, k


@pytest.mark.parametrize("lam_cm,invnu_fs,T", BATHS)
@pytest.mark.parametrize("K", [1, 2, 4])
def test_terminator_is_the_missing_zero_frequency_weight(oracle_mod, lam_cm, invnu_fs, T, K):
    lam, gamma = lam_cm * CM, 1 / (invnu_fs * 1e-15)
    bath = oracle_bath(lam, invnu_fs * 1e-15, T, K)
    assert delta_error(bath, lam, gamma, T) < 1e-10
    # a terminator that forgets one kept exponential, or counts S(0) whole, is caught
    bad = dict(bath, delta=bath["delta"] + (bath["c"][-1] / bath["nu"][-1]).real / O.HBAR)
    assert delta_error(bad, lam, gamma, T) > 1e-3
    bad = dict(bath, delta=bath["delta"] + spectral_zero_weight(lam, gamma, T) / O.HBAR)
    assert delta_error(bad, lam, gamma, T) > 1e-1


@pytest.mark.parametrize("lam_cm,invnu_fs,T", BATHS)
def test_circ_prefactor_is_the_imaginary_part(oracle_mod, lam_cm, invnu_fs, T):
    lam, gamma = lam_cm * CM, 1 / (invnu_fs * 1e-15)
    ts = np.linspace(5e-15, 200e-15, 20)
    bath = oracle_bath(lam, invnu_fs * 1e-15, T, 2)
    assert phi_error(bath, lam, gamma, ts) < 1e-8
    assert phi_error(dict(bath, phi=2 * bath["phi"]), lam, gamma, ts) > 0.5
    assert phi_error(dict(bath, phi=-bath["phi"]), lam, gamma, ts) > 1.0


# ---------------------------------------------------------------------------------
# linear absorption
# ---------------------------------------------------------------------------------

def write_la_config(path, H_cm, lam_cm, invnu_fs, K, T, depth, dt_fs, steps, observe, directions, strengths):
    """A linear-absorption config with one Drude bath per site."""
    n = len(H_cm)

    def mat(m):
        return "{" + ", ".join("{" + ", ".join(repr(float(v)) for v in row) + "}" for row in m) + "}"

    def vec(v):
        return "{" + ", ".join(repr(float(x)) for x in v) + "}"

    coupling = "{" + ", ".join("{%d}" % i for i in range(n)) + "}"
    text = f"""[program]
task=linear_absorption
observations={{(matrix_trace_static_fluorescence, out.dat)}}
observe_steps={observe}
[filtering]
strategy=none
first_layer=-1
[solver]
stepper_type=rk_rk4
step_size={dt_fs}e-15
steps={steps}
track_flows=false
flow_filename=
[system]
sites={n}
hamiltonian={mat(H_cm)}
ado_depth={depth}
[baths]
number={n}
max_per_site=1
coupling={coupling}
lambda={vec([lam_cm] * n)}
invnu={vec([invnu_fs] * n)}
Omega={vec([0] * n)}
matsubaras={K}
temperature={T}
[dipole]
directions={mat(directions)}
centers={mat(np.zeros((n, 3)))}
strengths={vec(strengths)}
tensor_prefactors={{1, 1, 1}}
tensor_components={{{{0, 0}}, {{1, 1}}, {{2, 2}}}}
"""
    with open(path, "w") as f:
        f.write(text)
    return O.parse_config(path)


def flipped_dipoles(cfg, t):
    dm, dp = PRISTINE["dipole_matrices"](cfg, t)
    return dm, -dp


def flipped_phase(e0, times):
    return np.exp(1j * e0 * times / O.HBAR)


PRISTINE = {"dipole_matrices": O.dipole_matrices, "rotating_frame_phase": O.rotating_frame_phase}
PLANTS = {"none": {}, "D+ sign": {"dipole_matrices": flipped_dipoles},
          "frame sign": {"rotating_frame_phase": flipped_phase}}


def fmo_system():
    fmo = O.parse_config(os.path.join(FMO_DIR, "fmo_linear_absorption_77K.cfg"))
    return fmo.hamiltonian.real / CM, fmo.dipole["directions"], fmo.dipole["strengths"]


def free_correlation(cfg):
    """mean_c mu_c^T exp(-i H t/hbar) mu_c, mu_c[i] = direction_i[c] * strength_i: the
    bath-free dipole correlation, from the config alone."""
    H = cfg.hamiltonian
    times = np.arange(cfg.steps // cfg.observe_steps + 1) * cfg.observe_steps * cfg.step_size
    out = np.zeros(len(times), dtype=np.complex128)
    comps = cfg.dipole["tensor_components"]
    for c, (a, b) in enumerate(comps):
        mu_a = cfg.dipole["directions"][:, a] * cfg.dipole["strengths"]
        mu_b = cfg.dipole["directions"][:, b] * cfg.dipole["strengths"]
        for r, t in enumerate(times):
            out[r] += cfg.dipole["tensor_prefactors"][c] * (mu_a @ scipy.linalg.expm(-1j * H * t / O.HBAR) @ mu_b)
    return times, out / len(comps)


@pytest.mark.parametrize("plant", list(PLANTS))
def test_absorption_without_bath_is_the_free_correlation(oracle_mod, tmp_path, monkeypatch, plant):
    for name, fn in PLANTS[plant].items():
        monkeypatch.setattr(O, name, fn)
    H, dirs, st = fmo_system()
    H = H + 11000 * np.eye(len(H))  # optical energies around 12,400 cm^-1
    cfg = write_la_config(str(tmp_path / "free.cfg"), H, 1e-9, 50.0, 2, 77.0, 2, 0.5, 400, 4, dirs, st)
    e0 = float(np.mean(np.diag(cfg.hamiltonian).real))
    times, C = O.linear_absorption(cfg, rotating_frame=e0)
    t_ref, ref = free_correlation(cfg)
    assert np.array_equal(times, t_ref)
    err = np.max(np.abs(C - ref)) / np.max(np.abs(ref))
    if plant == "none":
        assert err < 1e-7, err
    else:
        assert err > 0.1, err


def monomer_line_shape(cfg, times):
    """exp(-i eps t/hbar - g(t)) times the tensor-averaged mu^2, with the hierarchy's own
    c_k, nu_k, Delta, Phi (each pinned by quadrature above)."""
    K = cfg.baths_matsubaras
    bath = oracle_bath(cfg.baths_lambda[0], cfg.baths_invnu[0], cfg.baths_temperature, K)

    def F(v, t):
        return (np.exp(-v * t) + v * t - 1) / v ** 2

    g = np.array([np.sum(bath["c"] / O.HBAR * F(bath["nu"], t)) + bath["delta"] * t
                  - 1j * bath["phi"] * F(bath["nu"][0], t) for t in times])
    mu = cfg.dipole["directions"][0] * cfg.dipole["strengths"][0]
    mu2 = np.mean([mu[a] * mu[b] for a, b in cfg.dipole["tensor_components"]])
    eps = cfg.hamiltonian[0, 0].real
    return mu2 * np.exp(-1j * eps * times / O.HBAR - g)


@pytest.mark.parametrize("plant", list(PLANTS))
def test_monomer_absorption_is_the_cumulant_line_shape(oracle_mod, tmp_path, monkeypatch, plant):
    for name, fn in PLANTS[plant].items():
        monkeypatch.setattr(O, name, fn)
    errs = []
    for depth in (6, 10):
        cfg = write_la_config(str(tmp_path / "mono.cfg"), [[12000.0]], 35.0, 50.0, 3, 300.0, depth, 0.5, 800, 8,
                              [[0.6, 0.8, 0.0]], [2.0])
        e0 = cfg.hamiltonian[0, 0].real
        times, C = O.linear_absorption(cfg, rotating_frame=e0)
        exact = monomer_line_shape(cfg, times)
        errs.append(np.max(np.abs(C - exact)) / np.max(np.abs(exact)))
    if plant == "none":
        assert errs[1] < errs[0] / 100 and errs[1] < 1e-7, errs
    else:
        assert min(errs) > 0.1, errs


@pytest.mark.parametrize("plant", ["none", "frame sign"])
def test_rotating_frame_differs_by_rk4_truncation_only(oracle_mod, tmp_path, monkeypatch, plant):
    for name, fn in PLANTS[plant].items():
        monkeypatch.setattr(O, name, fn)
    H, dirs, st = fmo_system()
    diffs = []
    for dt in (0.5, 0.25, 0.125):
        n = int(round(100 / dt))
        cfg = write_la_config(str(tmp_path / "fmo.cfg"), H, 35.0, 50.0, 2, 77.0, 2, dt, n, n // 20, dirs, st)
        e0 = float(np.mean(np.diag(cfg.hamiltonian).real))
        _, rotating = O.linear_absorption(cfg, rotating_frame=e0)
        _, plain = O.linear_absorption(cfg)
        diffs.append(np.max(np.abs(rotating - plain)) / np.max(np.abs(plain)))
    ratios = [a / b for a, b in zip(diffs, diffs[1:])]
    if plant == "none":
        assert all(12 < r < 20 for r in ratios) and diffs[-1] < 1e-6, (diffs, ratios)
    else:
        assert min(diffs) > 0.1, diffs
