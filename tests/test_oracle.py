"""Pins the CPU oracle against what mathematics fixes.

The reference ships no tests or golden outputs, and it cannot be built here (its
submodules, Boost and an OpenCL runtime are absent), so each part of the oracle is
checked against an independent statement of the same thing: closed-form counts,
brute-force enumeration, a step-by-step restatement of the reference's partition
generator, a 50-digit evaluation of the prefactor formulas, the Bose function the Pade
table approximates, a dense operator-form Liouvillian, the no-bath (von Neumann)
limit, and the convergence order of RK4.  Each check is chosen so that a plausible
mistake (a dropped term, a wrong sign or index, a transposed operand) fails.
"""
import itertools
import math

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from helpers import FMO_DIR, config_path


# ---------------------------------------------------------------------------------
# partitions and the hierarchy graph
# ---------------------------------------------------------------------------------

def reference_partition_steps(n, s):
    """dm-heom/src/heom/integer_partition.cpp:41-105 restated step by step."""
    x = [1] * (n + 2)
    rows = []

    def emit(length):
        rows.append(x[1:length + 1] + [0] * (s - length))

    if n <= s:
        emit(n)
    if n > 1:
        x[0], x[1] = -1, 2
        h, m = 1, n - 1
        if m <= s:
            emit(m)
        while x[1] != n:
            if m - h > 1:
                h += 1
                x[h] = 2
                m -= 1
            else:
                j = m - 2
                while x[j] == x[m - 1]:
                    x[j] = 1
                    j -= 1
                h = j + 1
                x[h] = x[m - 1] + 1
                r = x[m] + x[m - 1] * (m - h - 1)
                x[m] = 1
                if m - h > 1:
                    x[m - 1] = 1
                m = h + r - 1
            if m <= s:
                emit(m)
    return rows


def brute_partitions(n, s):
    """All partitions of n into at most s parts (non-increasing), by brute force."""
    out = set()

    def rec(rem, mx, pre):
        if rem == 0:
            out.add(tuple(pre + [0] * (s - len(pre))))
            return
        if len(pre) == s:
            return
        for p in range(1, min(rem, mx) + 1):
            rec(rem - p, p, pre + [p])

    rec(n, n, [])
    return out


@pytest.mark.parametrize("n", range(1, 13))
@pytest.mark.parametrize("s", [1, 2, 3, 5, 8, 12, 16])
def test_partitions_match_reference_steps_and_brute_force(oracle_mod, n, s):
    got = O.partitions(n, s).tolist()
    assert got == reference_partition_steps(n, s)
    assert set(map(tuple, got)) == brute_partitions(n, s)
    assert len(got) == O.lib().oracle_partition_count(n, s)
    # increasing lexicographic order of the non-increasing part sequences
    assert all(got[r] < got[r + 1] for r in range(len(got) - 1))
    assert all(list(row) == sorted(row, reverse=True) for row in got)


@pytest.mark.parametrize("W,D", [(1, 1), (1, 5), (2, 3), (3, 3), (4, 2), (5, 4), (7, 4), (8, 3)])
def test_graph_structure(oracle_mod, W, D):
    g = O.build_graph(W, D)
    M = math.comb(W + D, W)
    assert g.nodes == M
    assert g.first_non_plus == math.comb(W + D - 1, W)
    assert g.nodes_per_depth.tolist() == [math.comb(d + W - 1, W - 1) for d in range(D + 1)]
    T = g.tuples.astype(int)
    # exactly the compositions with sum <= D, each once
    brute = {t for t in itertools.product(range(D + 1), repeat=W) if sum(t) <= D}
    assert set(map(tuple, T)) == brute and len(brute) == M
    # layer order; inside a layer, partition groups in generator order, each group's
    # permutations in increasing lexicographic order starting from the sorted tuple
    depth = T.sum(axis=1)
    assert np.all(np.diff(depth) >= 0)
    for d in range(1, D + 1):
        layer = [tuple(t) for t in T[depth == d]]
        parts = O.partitions(d, W).tolist()
        expect = []
        for p in parts:
            base = sorted(p)
            expect += sorted(set(itertools.permutations(base)))
        assert layer == expect
    # edges: plus/minus are unit steps, -2 where the step leaves the hierarchy
    index = {tuple(t): u for u, t in enumerate(T)}
    for u, t in enumerate(T):
        for m in range(W):
            up = list(t)
            up[m] += 1
            dn = list(t)
            dn[m] -= 1
            assert g.plus[u, m] == (index[tuple(up)] if sum(t) < D else -2)
            assert g.minus[u, m] == (index[tuple(dn)] if t[m] > 0 else -2)
            if g.plus[u, m] >= 0:
                assert g.minus[g.plus[u, m], m] == u


# ---------------------------------------------------------------------------------
# prefactors
# ---------------------------------------------------------------------------------

def mp_prefactors(cfg):
    """dm-heom/src/heom/instance.cpp:155-262 in 50-digit arithmetic (mpmath)."""
    import mpmath as mp

    mp.mp.dps = 50
    hbar = mp.mpf(O.HBAR)
    kb = mp.mpf(O.KB)
    beta = 1 / (kb * mp.mpf(cfg.baths_temperature))
    K = cfg.baths_matsubaras
    row = (K - 1) * 10
    eta_tab, xi_tab = oracle_pade()
    out = {"cbk": [], "pf_same": [], "pf_circ": [], "nu": []}
    for b in range(cfg.baths_number):
        lam = mp.mpf(cfg.baths_lambda[b])
        nu = mp.mpc(1 / mp.mpf(cfg.baths_invnu[b]), mp.mpf(cfg.baths_omega[b]))
        out["pf_circ"].append(nu * lam / hbar)
        delta = 2 * lam / (beta * hbar ** 2 * nu)
        for k in range(K):
            if k == 0:
                s = sum(2 * mp.mpf(eta_tab[row + j]) * nu ** 2 /
                        (mp.mpf(xi_tab[row + j]) ** 2 / (beta * hbar) ** 2 - nu ** 2) for j in range(K))
                c = 2 * lam / (beta * hbar) * (1 - s)
                nubk = nu
            else:
                eta, xi = mp.mpf(eta_tab[row + k]), mp.mpf(xi_tab[row + k])
                c = 4 * lam * nu / (beta * hbar) ** 2 * eta * xi / (xi ** 2 / (beta * hbar) ** 2 - nu ** 2)
                nubk = xi / (beta * hbar)
            out["cbk"].append(c)
            out["nu"].append(nubk)
            delta -= c / (hbar * nubk)
        out["pf_same"].append(delta)
    return out


def oracle_pade():
    """The Pade tables as the oracle holds them (read back through its C source)."""
    import re

    src = open(O.HERE + "/heom_oracle.c").read()

    def grab(name):
        body = src.split(f"static const long double {name}[100] = {{")[1].split("};")[0]
        return [v.strip().rstrip("L") for v in body.replace("\n", " ").split(",") if v.strip()]

    return grab("oracle_eta_pade"), grab("oracle_xi_pade")


def c2(arr):
    return arr[0::2] + 1j * arr[1::2]


@pytest.mark.parametrize("name", ["excitonic7_d4", "excitonic8_d8", "excitonic8_d6", "fmo"])
def test_prefactors_match_50_digit_formulas(oracle_mod, name):
    path = FMO_DIR + "/fmo_population_dynamics_77K.cfg" if name == "fmo" else config_path(name)
    cfg = O.parse_config(path)
    cfg.ado_depth = min(cfg.ado_depth, 3)
    inst = O.instance_from_config(cfg)
    ref = mp_prefactors(cfg)
    for key in ("cbk", "pf_same", "pf_circ"):
        got = c2(getattr(inst, key))
        want = np.array([complex(v) for v in ref[key]])
        assert np.max(np.abs(got - want) / np.abs(want)) < 2e-14, key
    assert np.allclose(inst.cbk_sqrt, np.sqrt(np.abs(c2(inst.cbk))), rtol=1e-15)
    # S_n = -sum_m n_m nu_m
    nu = np.array([complex(v) for v in ref["nu"]])
    T = inst.graph.tuples
    want = -(T * nu[None, :]).sum(axis=1)
    got = c2(inst.pf_same_sum)
    assert np.max(np.abs(got - want)) <= 1e-15 * np.max(np.abs(want))


@pytest.mark.parametrize("K", range(1, 11))
def test_pade_table_approximates_bose_function(K):
    """Row K-1 is a Pade spectrum decomposition of the Bose function,
    1/(1-e^{-x}) ~ 1/x + 1/2 + sum_j 2 eta_j x / (x^2 + xi_j^2), exact as x -> 0 and
    increasingly accurate with K; a wrong digit in a significant place breaks it."""
    eta_tab, xi_tab = oracle_pade()
    eta = np.array([float(v) for v in eta_tab[(K - 1) * 10:(K - 1) * 10 + K]])
    xi = np.array([float(v) for v in xi_tab[(K - 1) * 10:(K - 1) * 10 + K]])
    def err(xmax):
        x = np.linspace(0.05, xmax, 400)
        bose = 1.0 / (1.0 - np.exp(-x))
        approx = 1.0 / x + 0.5 + (2 * eta[None, :] * x[:, None] / (x[:, None] ** 2 + xi[None, :] ** 2)).sum(axis=1)
        return np.max(np.abs(approx - bose))

    # the accuracy of the [K-1/K] decomposition improves by orders of magnitude per K
    near = {1: 5e-4, 2: 1e-7, 3: 1e-11}.get(K, 1e-13)
    far = {1: 3e-2, 2: 3e-4, 3: 1e-6, 4: 1e-9, 5: 1e-12}.get(K, 1e-13)
    assert err(2.0) < near, (K, err(2.0))
    assert err(6.0) < far, (K, err(6.0))
    # poles are real, positive and distinct; the smallest approach the Matsubara 2 pi
    assert np.all(xi > 0) and len(set(xi)) == K
    if K >= 4:
        assert abs(np.min(xi) - 2 * np.pi) < 1e-6


def test_constants_are_the_reference_values(oracle_mod):
    assert O.HBAR == 1.05457148e-34
    assert O.KB == 1.380650239684840458953781387048428305608616777679979316e-23
    assert O.INVCM_TO_JOULE == 1.986456217327470903861498336971100097137908380913187891e-23
    # 1 cm^-1 = 2 pi hbar c / (1 cm); the reference's two constants come from different
    # CODATA releases and agree with each other only to about 5e-6
    assert abs(O.INVCM_TO_JOULE - 2 * np.pi * O.HBAR * 299792458.0 * 100) / O.INVCM_TO_JOULE < 1e-5


# ---------------------------------------------------------------------------------
# the RHS against a dense operator-form Liouvillian
# ---------------------------------------------------------------------------------

def small_instance(N, K, D, coupling, seed=0, omega=None, lam_scale=1.0, complex_h=False, T=300.0):
    rng = np.random.default_rng(seed)
    H = np.diag(rng.uniform(12000, 12600, N)).astype(np.complex128)
    for i in range(N):
        for j in range(i + 1, N):
            v = rng.uniform(-100, 100) + (1j * rng.uniform(-30, 30) if complex_h else 0)
            H[i, j], H[j, i] = v, np.conj(v)
    coupling = np.asarray(coupling, dtype=np.int32)
    B = int(coupling.max()) + 1
    lam = rng.uniform(20, 60, B) * lam_scale * O.INVCM_TO_JOULE
    invnu = rng.uniform(30, 120, B) * 1e-15
    om = (np.zeros(B) if omega is None else np.asarray(omega)) * (O.INVCM_TO_JOULE / O.HBAR)
    g = O.build_graph(B * K, D)
    inst = O.Instance(states=N, matsubaras=K, baths=B, coupling=coupling, hamiltonian=H * O.INVCM_TO_JOULE,
                      graph=g, lambda_=lam, invnu=invnu, omega=om, temperature=T)
    return inst, H * O.INVCM_TO_JOULE


def dense_liouvillian(inst, H):
    """The HEOM generator as a dense matrix, written with operators on the ADO space:
      d sigma_n = -(i/hbar)[H, sigma_n] + S_n sigma_n - sum_b Delta_b D_b(sigma_n)
                  + sum_m i sqrt((n_m+1)|c_m|/hbar) [V_b, sigma_{n+e_m}]
                  + sum_m sqrt(n_m hbar/|c_m|) (i c_m/hbar [V_b, .] + [k=0] Phi_b {V_b, .}) sigma_{n-e_m}
    with V_b = sum_s count_s(b) |s><s| and D_b(X)_ij = [i != j](count_i(b) + count_j(b)) X_ij.
    ADOs are indexed by brute-force enumeration, not by the oracle graph."""
    N, K, B = inst.states, inst.matsubaras, inst.baths
    W = inst.width
    D = inst.depth
    ados = [t for t in itertools.product(range(D + 1), repeat=W) if sum(t) <= D]
    idx = {t: a for a, t in enumerate(ados)}
    M = len(ados)
    dim = M * N * N
    I = np.eye(N)
    # superoperators acting on row-major vec(X): vec(A X B) = kron(A, B^T) vec(X)
    def left(A):
        return np.kron(A, I)

    def right(A):
        return np.kron(I, A.T)

    cnt = np.zeros((N, B))
    for s in range(N):
        for b in inst.coupling[s]:
            if b >= 0:
                cnt[s, b] += 1
    V = [np.diag(cnt[:, b]) for b in range(B)]
    comm = [left(v) - right(v) for v in V]
    anti = [left(v) + right(v) for v in V]
    cbk = c2(inst.cbk)
    delta = c2(inst.pf_same)
    phi = c2(inst.pf_circ)
    nu = -c2(inst.pf_same_sum)  # pf_same_sum of the unit tuples gives -nu_m
    unit = {m: idx[tuple(1 if k == m else 0 for k in range(W))] for m in range(W)}
    graph_index = {tuple(t): u for u, t in enumerate(inst.graph.tuples)}
    nu_m = np.array([nu[graph_index[tuple(1 if k == m else 0 for k in range(W))]] for m in range(W)])
    same = np.zeros((N * N, N * N), dtype=complex)
    offdiag = np.array([1.0 if i != j else 0.0 for i in range(N) for j in range(N)])
    for b in range(B):
        dvec = np.array([cnt[i, b] + cnt[j, b] for i in range(N) for j in range(N)]) * offdiag
        same -= delta[b] * np.diag(dvec)
    L = np.zeros((dim, dim), dtype=complex)
    com_h = -1j / O.HBAR * (left(H) - right(H))
    for a, t in enumerate(ados):
        blk = slice(a * N * N, (a + 1) * N * N)
        S = -sum(t[m] * nu_m[m] for m in range(W))
        L[blk, blk] += com_h + S * np.eye(N * N) + same
        for m in range(W):
            b, k = divmod(m, K)
            up = list(t)
            up[m] += 1
            if sum(up) <= D:
                c = idx[tuple(up)]
                L[blk, c * N * N:(c + 1) * N * N] += 1j * np.sqrt((t[m] + 1) * abs(cbk[m]) / O.HBAR) * comm[b]
            if t[m] > 0:
                dn = list(t)
                dn[m] -= 1
                c = idx[tuple(dn)]
                op = 1j * cbk[m] / O.HBAR * comm[b] + (phi[b] * anti[b] if k == 0 else 0)
                L[blk, c * N * N:(c + 1) * N * N] += np.sqrt(t[m] * O.HBAR / abs(cbk[m])) * op
    return L, ados, idx


def to_dense(inst, state, ados):
    n2 = inst.states ** 2
    s = state.reshape(inst.matrices, n2, 2)
    s = s[..., 0] + 1j * s[..., 1]
    gi = {tuple(t): u for u, t in enumerate(inst.graph.tuples)}
    return np.concatenate([s[gi[t]] for t in ados])


@pytest.mark.parametrize("case", [
    dict(N=3, K=1, D=3, coupling=[[0], [1], [2]]),
    dict(N=2, K=2, D=3, coupling=[[0], [1]]),
    dict(N=3, K=2, D=2, coupling=[[0], [1], [2]], omega=[0, 120, -80], complex_h=True),
    dict(N=3, K=1, D=2, coupling=[[-1], [0], [1]]),           # a ground state without bath
    dict(N=3, K=1, D=2, coupling=[[0, 1], [1, -1], [0, -1]]),  # several baths per state, shared
])
def test_rhs_matches_dense_liouvillian(oracle_mod, case):
    inst, H = small_instance(case["N"], case["K"], case["D"], case["coupling"], seed=3,
                             omega=case.get("omega"), complex_h=case.get("complex_h", False))
    L, ados, _ = dense_liouvillian(inst, H)
    rng = np.random.default_rng(9)
    y = rng.normal(size=inst.state_size)
    got = to_dense(inst, inst.rhs(y), ados)
    want = L @ to_dense(inst, y, ados)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_rhs_top_trace_is_conserved(oracle_mod):
    inst, _ = small_instance(4, 2, 3, [[0], [1], [2], [3]], seed=5)
    y = np.random.default_rng(2).normal(size=inst.state_size)
    f = inst.rhs(y)
    top = inst.top(f)
    scale = np.max(np.abs(inst.top(y))) * np.max(np.abs(inst.hamiltonian)) / O.HBAR
    assert abs(np.trace(top)) <= 1e-13 * scale


def test_no_bath_limit_is_von_neumann(oracle_mod):
    """With lambda -> 0 the top ADO follows sigma(t) = e^{-iHt/hbar} sigma e^{iHt/hbar}."""
    inst, H = small_instance(4, 1, 2, [[0], [1], [2], [3]], seed=7, lam_scale=1e-12)
    rho = np.zeros((4, 4), complex)
    rho[1, 1] = 1.0
    y = inst.state_with_top(rho)
    dt, n = 0.2e-15, 500
    tops = inst.propagate(y, dt, n, n)
    U = scipy.linalg.expm(-1j * H * n * dt / O.HBAR)
    want = U @ rho @ U.conj().T
    assert np.max(np.abs(tops[-1] - want)) < 1e-9


def test_rk4_is_fourth_order(oracle_mod):
    inst, H = small_instance(2, 1, 2, [[0], [1]], seed=11)
    L, ados, idx = dense_liouvillian(inst, H)
    rho = np.array([[0.7, 0.2 + 0.1j], [0.2 - 0.1j, 0.3]])
    y0 = inst.state_with_top(rho)
    T_end = 40e-15
    exact = scipy.linalg.expm(L * T_end) @ to_dense(inst, y0, ados)
    errs = []
    for n in (40, 80, 160):
        y = y0.copy()
        inst.propagate(y, T_end / n, n, n)
        errs.append(np.max(np.abs(to_dense(inst, y, ados) - exact)))
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    assert all(14.0 < r < 18.0 for r in ratios), (errs, ratios)


def test_rk_weighted_add_semantics(oracle_mod):
    rng = np.random.default_rng(0)
    import ctypes

    n = 37
    y = rng.normal(size=n)
    ks = [rng.normal(size=n) for _ in range(4)]
    c = np.array([1 / 6, 0.0, 1 / 3, 1 / 6])
    out = np.zeros(n)
    karr = (ctypes.c_void_p * 4)(*[k.ctypes.data for k in ks])
    O.lib().oracle_rk_weighted_add.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                               np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.float64),
                                               np.ctypeslib.ndpointer(np.float64), ctypes.c_void_p]
    O.lib().oracle_rk_weighted_add(n, 4, 0.5, out, y, c, ctypes.cast(karr, ctypes.c_void_p))
    want = y + 0.5 * (c[0] * ks[0] + c[2] * ks[2] + c[3] * ks[3])
    assert np.allclose(out, want, rtol=1e-15, atol=1e-15)


def test_norms_mmult_trace_are_textbook(oracle_mod):
    inst, _ = small_instance(3, 1, 2, [[0], [1], [2]], seed=1)
    rng = np.random.default_rng(4)
    y = rng.normal(size=inst.state_size)
    z = y.reshape(inst.matrices, 3, 3, 2)
    z = z[..., 0] + 1j * z[..., 1]
    assert np.allclose(inst.norms(y), np.sum(np.abs(z) ** 2, axis=(1, 2)), rtol=1e-14)
    mx = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    y2 = y.copy()
    inst.mmult_left(mx, y2)
    z2 = y2.reshape(inst.matrices, 3, 3, 2)
    assert np.allclose(z2[..., 0] + 1j * z2[..., 1], np.einsum("ik,nkj->nij", mx, z), rtol=1e-13, atol=1e-14)
    d = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    assert abs(O.trace(d, z[0]) - np.trace(d @ z[0])) < 1e-13
    assert abs(O.trace(None, z[0]) - np.trace(z[0])) < 1e-14


# ---------------------------------------------------------------------------------
# config grammar
# ---------------------------------------------------------------------------------

def test_fmo_example_parses(oracle_mod):
    cfg = O.parse_config(FMO_DIR + "/fmo_population_dynamics_77K.cfg")
    assert cfg.sites == 7 and cfg.ado_depth == 3 and cfg.baths_matsubaras == 2
    assert cfg.hamiltonian[0, 1] == pytest.approx(-87.7 * O.INVCM_TO_JOULE, rel=1e-15)
    assert cfg.baths_invnu[3] == pytest.approx(50e-15)
    assert [o[0] for o in cfg.observations] == ["matrix_complete", "matrix_diagonal", "matrix_trace_id"]
    la = O.parse_config(FMO_DIR + "/fmo_linear_absorption_77K.cfg")
    assert la.task == "linear_absorption" and la.dipole["directions"].shape == (7, 3)
    assert la.dipole["tensor_components"].tolist() == [[0, 0], [1, 1], [2, 2]]
