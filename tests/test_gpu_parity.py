"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Both sides receive the same instance arrays (reference ADO order, reference layout);
results come back in that order, so comparisons are keyed by ADO tuple.  Tolerances:
one RHS evaluation 1e-12 relative to the largest element; propagated populations and
ADO norms 1e-10 relative (norms relative to max(norm_n, 1e-12 max norm)).
"""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, instance, load_config, make_instance, max_abs_rel, random_state, rel_err

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


@pytest.fixture(params=["pipelined", "tile", "split"], autouse=True)
def stage_kernel(request, monkeypatch):
    """Every test runs with the three stage-kernel configurations for N <= 8: the
    persistent pipelined kernel (where it applies: N >= 6, few modes), the
    one-workgroup-per-tile kernel, and the tile kernel with each tile split over
    workgroups (the default for small hierarchies; where N allows)."""
    monkeypatch.setenv("HEOM_STAGE_KERNEL", request.param)
    return request.param


def expected_kernel(inst, requested, complex_bytes=16):
    """The pipelined kernel holds two state tiles (N rows of ceil(N/4)*4 element slots x
    64 ADOs) and two (W+2)-row table tiles (rounded to an even row count) in 160 KiB of
    LDS with 4 KiB spare; with complex64 state it runs at N = 8 only."""
    N, W = inst.states, inst.width
    rs = (N + 3) // 4 * 4
    cap = min(64, (163840 - 2 * N * rs * 64 * complex_bytes) // 1024 - 4) & ~1
    shape = N >= 6 if complex_bytes == 16 else N == 8
    ok = requested == "pipelined" and shape and ((W + 2) & ~1) <= cap
    return "pipelined" if ok else "tile"


def gpu_rhs(inst, y, mask=None, kernel=None):
    with heom_amd.Handle.from_instance(inst) as h:
        if kernel is not None:
            assert h.stage_kernel() == expected_kernel(inst, kernel)
        if mask is not None:
            h.set_mask(mask)
        h.set_state(y)
        return h.rhs()


def check_rhs(inst, seed=1, mask=None, kernel=None):
    y = random_state(inst, seed)
    if mask is not None:
        inst.mask[:] = mask
    ref = inst.rhs(y)
    got = gpu_rhs(inst, y, mask, kernel)
    inst.mask[:] = O.MASK_TRUE
    err = max_abs_rel(got, ref)
    print(f"rhs N={inst.states} M={inst.matrices} W={inst.width}: max rel err {err:.3e}", flush=True)
    assert err < 1e-12
    # per-ADO: every ADO's derivative matches relative to its own size
    a, b = as_complex(got, inst), as_complex(ref, inst)
    per = np.max(np.abs(a - b), axis=(1, 2)) / np.maximum(np.max(np.abs(b), axis=(1, 2)), 1e-30)
    big = np.max(np.abs(b), axis=(1, 2)) > 1e-6 * np.max(np.abs(b))
    assert np.max(per[big]) < 1e-10


@pytest.mark.parametrize("name,depth", [("excitonic7_d4", None), ("excitonic8_d8", 3), ("excitonic8_d6", None)])
def test_rhs_workloads(oracle_mod, stage_kernel, name, depth):
    _, inst = instance(name, depth=depth)
    check_rhs(inst, kernel=stage_kernel)


def test_rhs_fmo_example(oracle_mod, stage_kernel):
    cfg = O.parse_config(__import__("helpers").FMO_DIR + "/fmo_population_dynamics_77K.cfg")
    inst = O.instance_from_config(cfg)
    check_rhs(inst, seed=3, kernel=stage_kernel)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
def test_rhs_all_state_counts(oracle_mod, stage_kernel, N):
    inst = make_instance(N, N, 2, 3, [[i] for i in range(N)], seed=N)
    check_rhs(inst, seed=N, kernel=stage_kernel)


def test_rhs_complex_hamiltonian_and_omega(oracle_mod):
    inst = make_instance(5, 5, 2, 3, [[i] for i in range(5)], seed=7, complex_h=True,
                         omega=[0, 150, -150, 80, 0])
    check_rhs(inst, seed=7)


def test_rhs_general_coupling(oracle_mod):
    # two baths per state, shared baths, missing entries (-1), a state without bath
    coupling = [[0, 1], [1, -1], [2, 0], [-1, -1], [3, 2]]
    inst = make_instance(5, 4, 2, 3, coupling, seed=11)
    check_rhs(inst, seed=11)


def test_rhs_with_mask(oracle_mod):
    _, inst = instance("excitonic7_d4")
    rng = np.random.default_rng(5)
    mask = rng.choice([0, 1, 2], size=inst.matrices, p=[0.3, 0.5, 0.2]).astype(np.uint8)
    mask[0] = 1
    check_rhs(inst, seed=5, mask=mask)


def test_state_roundtrip_and_top(oracle_mod):
    _, inst = instance("excitonic7_d4")
    y = random_state(inst, 9)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        assert np.array_equal(h.get_state(), y)
        top = h.get_top()
        assert np.array_equal(top, inst.top(y))
        new_top = np.arange(49).reshape(7, 7) * (1 + 0.5j)
        h.set_top(new_top)
        s = h.get_state()
        assert np.array_equal(inst.top(s), new_top)
        assert np.array_equal(s[98:], y[98:])


def test_norms_and_threshold_mask(oracle_mod):
    _, inst = instance("excitonic7_d4")
    y = random_state(inst, 13)
    ref = inst.norms(y)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        got = h.norms()
        assert rel_err(got, ref) < 1e-13
        assert np.all(h.get_mask() == 1)
        # threshold deactivation (cl/heom/hierarchy_norm.cl:41-44), PROTECTED survives
        mask = np.ones(inst.matrices, np.uint8)
        mask[5] = 2
        h.set_mask(mask)
        thr = float(np.median(ref))
        h.norms(thr)
        m = h.get_mask()
        expect = np.where((ref - thr <= 0) & (mask != 2), 0, mask)
        assert np.array_equal(m, expect)
        # the RHS now drops terms that read masked neighbours
        got_rhs = h.rhs()
    inst.mask[:] = expect
    ref_rhs = inst.rhs(y)
    inst.mask[:] = 1
    assert max_abs_rel(got_rhs, ref_rhs) < 1e-12


def test_mmult_left(oracle_mod):
    _, inst = instance("excitonic7_d4")
    y = random_state(inst, 17)
    rng = np.random.default_rng(17)
    mx = rng.normal(size=(7, 7)) + 1j * rng.normal(size=(7, 7))
    ref = y.copy()
    inst.mmult_left(mx, ref)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        h.mmult_left(mx)
        got = h.get_state()
    assert max_abs_rel(got, ref) < 1e-14


def test_rk4_populations_excitonic7(oracle_mod, stage_kernel):
    """The full 7-site depth-4 run: 1000 RK4 steps of 1 fs, observed every 10 steps."""
    cfg, inst = instance("excitonic7_d4")
    y0 = inst.state_with_top(cfg.rho_init)
    y = y0.copy()
    ref_tops = inst.propagate(y, cfg.step_size, cfg.steps, cfg.observe_steps)
    with heom_amd.Handle.from_instance(inst) as h:
        assert h.stage_kernel() == expected_kernel(inst, stage_kernel)
        h.set_state(y0)
        tops = h.propagate(cfg.step_size, cfg.steps, cfg.observe_steps)
        final = h.get_state()
        norms = h.norms()
    pops = np.real(np.diagonal(tops, axis1=1, axis2=2))
    ref_pops = np.real(np.diagonal(ref_tops, axis1=1, axis2=2))
    err = rel_err(pops, ref_pops, floor_frac=1e-12)
    print(f"populations max rel err {err:.3e}", flush=True)
    assert err < 1e-10
    ref_norms = inst.norms(y)
    nerr = rel_err(norms, ref_norms, floor_frac=1e-12)
    print(f"ADO norms max rel err {nerr:.3e}", flush=True)
    assert nerr < 1e-10
    assert max_abs_rel(final, y) < 1e-10


def test_step_rk4_matches_propagate(oracle_mod):
    cfg, inst = instance("excitonic7_d4")
    y0 = inst.state_with_top(cfg.rho_init)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y0)
        tops = h.propagate(cfg.step_size, 40, 20)
        a = h.get_state()
        h.set_state(y0)
        h.step_rk4(cfg.step_size, 20)
        t1 = h.get_top()
        h.step_rk4(cfg.step_size, 20)
        h.sync()
        b = h.get_state()
        t = h.timing()
    assert np.array_equal(a, b)
    assert np.array_equal(tops[1], t1)
    assert t.rk4_steps >= 80 and t.rk4_seconds > 0


def test_excitonic8_full_size_one_step(oracle_mod, stage_kernel):
    """735,471 ADOs: one RK4 step against the oracle, then invariants after 10 more."""
    cfg, inst = instance("excitonic8_d8")
    y = random_state(inst, 21)
    ref = y.copy()
    inst.rk4_step(ref, cfg.step_size)
    with heom_amd.Handle.from_instance(inst) as h:
        assert h.stage_kernel() == expected_kernel(inst, stage_kernel)
        h.set_state(y)
        h.step_rk4(cfg.step_size, 1)
        got = h.get_state()
        assert max_abs_rel(got, ref) < 1e-12
        a, b = as_complex(got, inst), as_complex(ref, inst)
        na, nb = np.sum(np.abs(a) ** 2, axis=(1, 2)), np.sum(np.abs(b) ** 2, axis=(1, 2))
        assert rel_err(na, nb, floor_frac=1e-12) < 1e-10
        # physical start: trace conserved, top stays Hermitian
        h.set_state(inst.state_with_top(cfg.rho_init))
        h.step_rk4(cfg.step_size, 10)
        top = h.get_top()
    assert abs(np.trace(top) - 1.0) < 1e-12
    assert np.max(np.abs(top - top.conj().T)) < 1e-12


def test_errors_are_reported(oracle_mod):
    _, inst = instance("excitonic7_d4")
    with heom_amd.Handle.from_instance(inst) as h:
        with pytest.raises(heom_amd.HeomError):
            h.step_rk4(-1.0, 1)
        with pytest.raises(ValueError):
            h.set_state(np.zeros(3))
        with pytest.raises(heom_amd.HeomError):
            h.set_mask(np.full(inst.matrices, 7, np.uint8))


@pytest.mark.parametrize("D", [1, 2, 62])
def test_rhs_single_mode_edge_depths(oracle_mod, stage_kernel, D):
    """One bath mode (W = 1): M = D + 1 ADOs, down to 2; depth 62 reaches the largest
    tuple entry the kernels take (their sqrt(n_m + 1) table has 64 entries).  State 1
    couples to no bath."""
    inst = make_instance(3, 1, 1, D, [[0], [-1], [0]], seed=D)
    assert inst.matrices == D + 1
    check_rhs(inst, seed=D, kernel=stage_kernel)
    y = random_state(inst, D)
    ref = inst.propagate(y.copy(), 1e-15, 20, 5)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        tops = h.propagate(1e-15, 20, 5)
    assert max_abs_rel(tops, ref) < 1e-10


def test_tuple_entry_above_62_is_refused(oracle_mod):
    inst = make_instance(2, 1, 1, 63, [[0], [0]], seed=1)
    with pytest.raises(heom_amd.HeomError, match="62"):
        heom_amd.Handle.from_instance(inst)
