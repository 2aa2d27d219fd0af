"""FP32 storage (HEOM_PRECISION_FP32): complex64 ADO planes and FP32 RHS, with the top
ADO and the tier-1 ADOs re-evaluated and accumulated in FP64.

The bar is the population error against the FP64 path: at most 1e-6 absolute over
1000 fs.  One FP32 RHS evaluation is checked against the FP64 oracle at FP32
resolution (1e-5 relative to the largest element), and the tier-0/1 ADOs must
round-trip exactly.
"""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
FP32 = heom_amd.HEOM_PRECISION_FP32


@pytest.fixture(params=["pipelined", "tile", "split"], autouse=True)
def stage_kernel(request, monkeypatch):
    monkeypatch.setenv("HEOM_STAGE_KERNEL", request.param)
    return request.param


def tier01(inst):
    return np.nonzero(inst.graph.tuples.sum(axis=1) <= 1)[0]


def test_fp32_state_roundtrip(oracle_mod):
    _, inst = instance("excitonic7_d4")
    y = random_state(inst, 3)
    with heom_amd.Handle.from_instance(inst, precision=FP32) as h:
        h.set_state(y)
        got = h.get_state()
        top = h.get_top()
    a, b = as_complex(got, inst), as_complex(y, inst)
    t = tier01(inst)
    assert np.array_equal(a[t], b[t])  # FP64 shadows
    rest = np.setdiff1d(np.arange(inst.matrices), t)
    assert np.array_equal(a[rest], b[rest].astype(np.complex64).astype(np.complex128))
    assert np.array_equal(top, b[0])


@pytest.mark.parametrize("name,depth", [("excitonic7_d4", None), ("excitonic8_d8", 3)])
def test_fp32_rhs(oracle_mod, name, depth):
    _, inst = instance(name, depth=depth)
    y = random_state(inst, 7)
    y32 = y.astype(np.float32).astype(np.float64)  # the values the FP32 planes hold
    t = tier01(inst)
    ys = as_complex(y32, inst)
    ys[t] = as_complex(y, inst)[t]
    y_in = np.ascontiguousarray(np.stack([ys.real, ys.imag], -1).reshape(-1))
    ref = inst.rhs(y_in)
    with heom_amd.Handle.from_instance(inst, precision=FP32) as h:
        h.set_state(y)
        got = h.rhs()
    err = max_abs_rel(got, ref)
    print(f"FP32 rhs {name}: max rel deviation {err:.2e}", flush=True)
    assert err < 1e-5
    # the tier-0/1 rows are evaluated in FP64
    a, b = as_complex(got, inst), as_complex(ref, inst)
    assert np.max(np.abs(a[t] - b[t])) / np.max(np.abs(b[t])) < 1e-12


def test_fp32_populations_1000fs(oracle_mod):
    """The 7-site depth-4 workload: 1000 steps of 1 fs, FP32 storage against FP64."""
    cfg, inst = instance("excitonic7_d4")
    y0 = inst.state_with_top(cfg.rho_init)
    pops = {}
    for prec in (heom_amd.HEOM_PRECISION_FP64, FP32):
        with heom_amd.Handle.from_instance(inst, precision=prec) as h:
            h.set_state(y0)
            tops = h.propagate(cfg.step_size, cfg.steps, cfg.observe_steps)
        pops[prec] = np.real(np.diagonal(tops, axis1=1, axis2=2))
        tr = np.real(np.trace(tops, axis1=1, axis2=2))
        assert np.max(np.abs(tr - 1.0)) < 1e-6
    err = np.max(np.abs(pops[FP32] - pops[heom_amd.HEOM_PRECISION_FP64]))
    print(f"FP32 storage: max population error over 1000 fs {err:.2e}", flush=True)
    assert err < 1e-6


def test_fp32_batch_matches_single(oracle_mod):
    """FP32 batches: bit-identical to FP32 single-instance handles."""
    _, a = instance("excitonic7_d4", depth=3)
    _, b = instance("excitonic7_d4", depth=4)
    top = np.zeros((7, 7), np.complex128)
    top[1, 1] = 1.0
    with heom_amd.Handle.from_instances([a, b], precision=FP32) as hb:
        hb.clear_state()
        hb.set_top(top)
        rec = hb.propagate(1e-15, 40, 10)
    for k, inst in enumerate((a, b)):
        with heom_amd.Handle.from_instance(inst, precision=FP32) as h:
            h.clear_state()
            h.set_top(top)
            rec1 = h.propagate(1e-15, 40, 10)
        assert np.array_equal(rec[:, k], rec1), k


# (states, modes per bath, complex Hamiltonian): site-local baths, so a paired plane
# couples to 2K modes (term slots 2, 4, 8); every N of the paired kernel, both
# Hamiltonian forms
PAIR_CASES = [(4, 1, False), (5, 2, True), (6, 4, False), (7, 1, True), (7, 3, False), (8, 2, False),
              (8, 4, True)]


@pytest.mark.parametrize("N,K,complex_h", PAIR_CASES)
def test_fp32_paired_planes(oracle_mod, stage_kernel, N, K, complex_h):
    """The paired complex64 layout (elements (i, j) and (j, i) in one 16-byte plane slot)
    runs whenever the pipelined kernel is asked for with 4 to 8 states; one RHS and ten
    RK4 steps against the FP64 oracle."""
    from test_gpu_parity import make_instance

    inst = make_instance(N, N, K, 3, [[i] for i in range(N)], seed=10 * N + K, complex_h=complex_h)
    y = random_state(inst, N + K)
    t = tier01(inst)
    y32 = as_complex(y.astype(np.float32).astype(np.float64), inst)
    y32[t] = as_complex(y, inst)[t]
    y_in = np.ascontiguousarray(np.stack([y32.real, y32.imag], -1).reshape(-1))
    ref = inst.rhs(y_in)
    dt, steps = 1e-15, 10
    ys = y_in.copy()
    for _ in range(steps):
        inst.rk4_step(ys, dt, np.empty(5 * inst.state_size))
    with heom_amd.Handle.from_instance(inst, precision=FP32) as h:
        assert h.stage_kernel() == ("pipelined" if stage_kernel == "pipelined" else "tile")
        h.set_state(y)
        got = h.rhs()
        h.step_rk4(dt, steps)
        stepped = h.get_state()
    e1, e2 = max_abs_rel(got, ref), max_abs_rel(stepped, ys)
    print(f"N={N} K={K} complex_h={complex_h}: rhs {e1:.1e}, {steps} steps {e2:.1e}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5
