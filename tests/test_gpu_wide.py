"""Large systems (8 < N <= 32 states, stage_wide.hip): the 24-site workload against the
oracle, in FP64 (1e-10 relative) and FP32 storage, N = 9, 12, 16 and 32 in FP64 (both
LDS instances of the kernel), every workgroup size the autotuner tries, and the
FP32-storage population error against FP64 over 1000 fs at depth 4 (20,475 complex64
24 x 24 ADOs)."""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
FP32 = heom_amd.HEOM_PRECISION_FP32


def tier01(inst):
    return np.nonzero(inst.graph.tuples.sum(axis=1) <= 1)[0]


@pytest.mark.parametrize("depth", [1, 2])
def test_wide_rhs_fp64_matches_oracle(oracle_mod, depth):
    _, inst = instance("excitonic24_d4", depth=depth)
    y = random_state(inst, 3)
    ref = inst.rhs(y)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        assert np.array_equal(h.get_state(), y)
        got = h.rhs()
    err = max_abs_rel(got, ref)
    print(f"N=24 depth {depth} ({inst.matrices} ADOs): FP64 rhs max rel deviation {err:.2e}", flush=True)
    assert err < 1e-12


def fp32_case(N):
    if N == 24:
        return instance("excitonic24_d4", depth=2)[1]
    from test_gpu_parity import make_instance
    return make_instance(N, N, 1, 2, [[i] for i in range(N)], seed=N)


@pytest.mark.parametrize("N", [12, 24, 32])
def test_wide_rhs_fp32(oracle_mod, N):
    """FP32 storage: the RHS against the oracle on the FP32-rounded state (tier-0/1 ADOs
    kept in FP64), for both LDS instances (N = 12 and 24: 48 KB, N = 32: 144 KB)."""
    inst = fp32_case(N)
    y = random_state(inst, 7)
    t = tier01(inst)
    ys = as_complex(y.astype(np.float32).astype(np.float64), inst)
    ys[t] = as_complex(y, inst)[t]
    ref = inst.rhs(np.ascontiguousarray(np.stack([ys.real, ys.imag], -1).reshape(-1)))
    with heom_amd.Handle.from_instance(inst, precision=FP32) as h:
        h.set_state(y)
        got = h.rhs()
    err = max_abs_rel(got, ref)
    print(f"N={N} FP32 rhs max rel deviation {err:.2e}", flush=True)
    assert err < 1e-5
    a, b = as_complex(got, inst), as_complex(ref, inst)
    assert np.max(np.abs(a[t] - b[t])) / np.max(np.abs(b[t])) < 1e-12


def test_wide_propagation_matches_oracle(oracle_mod):
    cfg, inst = instance("excitonic24_d4", depth=2)
    y = inst.state_with_top(cfg.rho_init)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y)
        tops = h.propagate(cfg.step_size, 60, 10)
        h.set_state(y)
        pops, tr = h.propagate_observe(cfg.step_size, 60, 10, np.eye(inst.states))
    ref = inst.propagate(y.copy(), cfg.step_size, 60, 10)
    err = max_abs_rel(tops, ref)
    print(f"N=24 60 steps: max rel deviation {err:.2e}", flush=True)
    assert err < 1e-10
    assert np.allclose(pops[:, 0], np.real(np.diagonal(ref, axis1=1, axis2=2)), rtol=0, atol=1e-12)
    assert np.allclose(tr[:, 0], np.trace(ref, axis1=1, axis2=2), rtol=0, atol=1e-12)


def test_wide_batch_and_mmult(oracle_mod):
    _, a = instance("excitonic24_d4", depth=1)
    _, b = instance("excitonic24_d4", depth=2)
    N = a.states
    rng = np.random.default_rng(4)
    mx = rng.normal(size=(2, N, N)) + 1j * rng.normal(size=(2, N, N))
    ys = [random_state(a, 1), random_state(b, 2)]
    with heom_amd.Handle.from_instances([a, b]) as hb:
        hb.set_state(np.concatenate(ys))
        hb.step_rk4(1e-15, 5)
        sb = hb.get_state()
        hb.mmult_left(mx)
        mb = hb.get_state()
    off = 0
    for k, inst in enumerate((a, b)):
        with heom_amd.Handle.from_instance(inst) as h:
            h.set_state(ys[k])
            h.step_rk4(1e-15, 5)
            s1 = h.get_state()
        n = inst.state_size
        assert np.array_equal(sb[off:off + n], s1), k
        want = s1.copy()
        inst.mmult_left(mx[k], want)
        assert max_abs_rel(mb[off:off + n], want) < 1e-13
        off += n


def test_wide_fp32_population_error_1000fs(oracle_mod):
    """The 24-site workload at depth 4: FP32 storage against FP64, 1000 steps of 1 fs."""
    cfg, inst = instance("excitonic24_d4")
    assert inst.matrices == 20475
    y0 = inst.state_with_top(cfg.rho_init)
    pops = {}
    for prec in (heom_amd.HEOM_PRECISION_FP64, FP32):
        with heom_amd.Handle.from_instance(inst, precision=prec) as h:
            h.set_state(y0)
            tops = h.propagate(cfg.step_size, cfg.steps, cfg.observe_steps)
        pops[prec] = np.real(np.diagonal(tops, axis1=1, axis2=2))
        assert np.max(np.abs(np.real(np.trace(tops, axis1=1, axis2=2)) - 1.0)) < 1e-6
    err = np.max(np.abs(pops[FP32] - pops[heom_amd.HEOM_PRECISION_FP64]))
    print(f"24 sites, depth 4: FP32-storage max population error over 1000 fs {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("N", [9, 12, 16, 32])
def test_wide_state_counts_match_oracle(oracle_mod, N):
    """RHS and 5 RK4 steps for N = 9..32 (one bath mode per site, depth 2): up to N = 17
    an 8-ADO FP64 tile fits the 48 KB LDS instance of the kernel, N = 32 runs the 144 KB
    one."""
    from test_gpu_parity import make_instance
    inst = make_instance(N, N, 1, 2, [[i] for i in range(N)], seed=N)
    y = random_state(inst, N)
    ref_rhs = inst.rhs(y)
    ref = inst.propagate(y.copy(), 1e-15, 5, 5)
    with heom_amd.Handle.from_instance(inst) as h:
        assert h.stage_kernel() == "wide"
        h.set_state(y)
        got_rhs = h.rhs()
        tops = h.propagate(1e-15, 5, 5)
    e1, e2 = max_abs_rel(got_rhs, ref_rhs), max_abs_rel(tops, ref)
    print(f"wide N={N} ({inst.matrices} ADOs): rhs {e1:.1e}, 5 steps {e2:.1e}", flush=True)
    assert e1 < 1e-12 and e2 < 1e-10


@pytest.mark.parametrize("prec", [heom_amd.HEOM_PRECISION_FP64, FP32])
def test_wide_every_tile_size_after_autotune(oracle_mod, prec):
    """The autotuner times 8-, 4-, 2- and 1-ADO workgroups on the 24-site hierarchy (both
    LDS instances) and restores the state; the tuned handle then matches an untuned one."""
    _, inst = instance("excitonic24_d4", depth=2)
    y = random_state(inst, 9)
    with heom_amd.Handle.from_instance(inst, precision=prec) as h:
        h.set_state(y)
        before = h.get_state()  # FP32 storage rounds y
        res = h.autotune(steps=1)
        assert {r["ados_per_workgroup"] for r in res} >= {8, 4, 2, 1}
        assert np.array_equal(h.get_state(), before)
        h.set_state(y)
        h.step_rk4(1e-15, 3)
        tuned = h.get_state()
    with heom_amd.Handle.from_instance(inst, precision=prec) as h:
        h.set_state(y)
        h.step_rk4(1e-15, 3)
        plain = h.get_state()
    assert np.array_equal(tuned, plain)


@pytest.mark.parametrize("prec", [heom_amd.HEOM_PRECISION_FP64, FP32])
def test_wide_forced_tile_sizes_are_bitwise_equal(prec):
    """Every tile size the autotuner offers, forced with set_stage_config on the 24-site
    hierarchy at depth 2 (325 ADOs in 384 slots: 1- and 2-ADO tiles give 384 and 192 workgroups, so
    the XCD-aware workgroup order is active), gives bitwise the 8-ADO result: the tile
    size changes only which workgroup computes an ADO."""
    _, inst = instance("excitonic24_d4", depth=2)
    y = random_state(inst, 11)
    results = {}
    with heom_amd.Handle.from_instance(inst, precision=prec) as h:
        configs = h.autotune(steps=1)
    for cfg in configs:
        with heom_amd.Handle.from_instance(inst, precision=prec) as h:
            h.set_stage_config(cfg)
            h.set_state(y)
            h.step_rk4(1e-15, 3)
            results[cfg["ados_per_workgroup"]] = h.get_state()
    assert set(results) >= {8, 4, 2, 1}
    for T, got in results.items():
        assert np.array_equal(got, results[8]), T


def test_set_stage_config_refuses_invalid():
    _, inst = instance("excitonic24_d4", depth=2)
    with heom_amd.Handle.from_instance(inst) as h:
        with pytest.raises(heom_amd.HeomError):
            h.set_stage_config({"kernel": "wide", "xcd_chunk": 1, "workgroups": 0, "ados_per_workgroup": 3})
        with pytest.raises(heom_amd.HeomError):
            h.set_stage_config({"kernel": "pipelined", "xcd_chunk": 1, "workgroups": 256, "ados_per_workgroup": 64})
