"""Batches whose instances differ in bath prefactors (lambda, gamma), coupling, number of
modes and number of states, stepped in one handle: each instance against the oracle run
of that instance alone (1e-12 relative) and against a single-instance handle (the
padded states and modes only add exact zeros, so 1e-14), in FP64 and FP32 storage; the
padded rows and columns stay exactly zero; the adjoint refuses such a batch."""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, instance, max_abs_rel
from test_gpu_parity import make_instance

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
FP32 = heom_amd.HEOM_PRECISION_FP32


def bath_variant(name, depth, lam_cm, invnu_fs, mode="identity"):
    """A config's instance with its lambda and 1/gamma replaced (every bath)."""
    cfg, _ = instance(name, depth=depth)
    cfg.baths_lambda = np.full(len(cfg.baths_lambda), lam_cm * O.INVCM_TO_JOULE)
    cfg.baths_invnu = np.full(len(cfg.baths_invnu), invnu_fs * 1e-15)
    return cfg, O.instance_from_config(cfg, mode)


def pad_state(insts, states, N):
    """Per-instance reference-layout states -> one batch state with N x N ADOs."""
    parts = []
    for inst, y in zip(insts, states):
        z = as_complex(y, inst)
        p = np.zeros((inst.matrices, N, N), dtype=np.complex128)
        p[:, :inst.states, :inst.states] = z
        parts.append(p)
    z = np.concatenate(parts)
    return np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1))


def split_state(y, insts, N):
    z = y.reshape(-1, N, N, 2)
    z = z[..., 0] + 1j * z[..., 1]
    out, off = [], 0
    for inst in insts:
        out.append(z[off:off + inst.matrices])
        off += inst.matrices
    return out


def mixed_instances():
    """7 sites with three (lambda, gamma) settings, an 8-state system with a bath-free
    ground state (the absorption shape), and 5 sites with two modes per bath (W = 10)."""
    out = []
    for lam, invnu in ((35.0, 50.0), (60.0, 100.0), (20.0, 25.0)):
        out.append(bath_variant("excitonic7_d4", 3, lam, invnu)[1])
    out.append(make_instance(8, 7, 1, 3, [[-1]] + [[i] for i in range(7)], seed=5))
    out.append(make_instance(5, 5, 2, 3, [[i] for i in range(5)], seed=6))
    return out


def random_states(insts, seed):
    rng = np.random.default_rng(seed)
    ys = []
    for inst in insts:
        depth = inst.graph.tuples.sum(axis=1)
        n = inst.states
        z = (rng.normal(size=(inst.matrices, n, n)) + 1j * rng.normal(size=(inst.matrices, n, n)))
        z *= (0.3 ** depth)[:, None, None]
        z[0] = z[0] + z[0].conj().T
        ys.append(np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1)))
    return ys


@pytest.mark.parametrize("prec", [heom_amd.HEOM_PRECISION_FP64, FP32])
def test_mixed_batch_matches_single_instances(oracle_mod, prec):
    insts = mixed_instances()
    ys = random_states(insts, 7)
    N = max(i.states for i in insts)
    dt, steps = 1e-15, 20
    with heom_amd.Handle.from_instances(insts, precision=prec) as h:
        assert h.states == N and h.count == len(insts)
        h.set_state(pad_state(insts, ys, N))
        rhs = split_state(h.rhs(), insts, N)
        h.step_rk4(dt, steps)
        got = split_state(h.get_state(), insts, N)
        pops, _ = h.propagate_observe(dt, 2, 1)
    for k, inst in enumerate(insts):
        n = inst.states
        # the padded rows and columns stay exactly zero
        assert not np.any(got[k][:, n:, :]) and not np.any(got[k][:, :, n:])
        assert not np.any(pops[:, k, n:])
        with heom_amd.Handle.from_instance(inst, precision=prec) as s:
            s.set_state(ys[k])
            r1 = as_complex(s.rhs(), inst)
            s.step_rk4(dt, steps)
            single = as_complex(s.get_state(), inst)
        tol = 1e-14 if prec != FP32 else 1e-6
        assert max_abs_rel(rhs[k][:, :n, :n], r1) < tol, k
        assert max_abs_rel(got[k][:, :n, :n], single) < tol, k
        if prec != FP32:
            y = ys[k].copy()
            ref_rhs = as_complex(inst.rhs(y), inst)
            for _ in range(steps):
                inst.rk4_step(y, dt, np.empty(5 * inst.state_size))
            e1 = max_abs_rel(rhs[k][:, :n, :n], ref_rhs)
            e2 = max_abs_rel(got[k][:, :n, :n], as_complex(y, inst))
            print(f"instance {k} (N={n}, W={inst.width}): rhs {e1:.1e}, {steps} steps {e2:.1e}", flush=True)
            assert e1 < 1e-12 and e2 < 1e-12


def test_mixed_batch_top_and_observables():
    """Per-instance tops and operators of their own sizes are padded by the binding."""
    insts = mixed_instances()
    N = max(i.states for i in insts)
    tops = []
    for inst in insts:
        t = np.zeros((inst.states, inst.states), dtype=np.complex128)
        t[0, 0] = 1.0
        tops.append(t)
    ops = [np.diag(np.arange(1.0, inst.states + 1)) for inst in insts]
    with heom_amd.Handle.from_instances(insts) as h:
        h.set_top(tops)
        pops, tr = h.propagate_observe(1e-15, 10, 5, ops)
    for k, inst in enumerate(insts):
        with heom_amd.Handle.from_instance(inst) as s:
            s.set_top(tops[k])
            p1, t1 = s.propagate_observe(1e-15, 10, 5, ops[k])
        assert max_abs_rel(pops[:, k, :inst.states], p1[:, 0]) < 1e-14
        assert max_abs_rel(tr[:, k], t1[:, 0]) < 1e-14


def test_adjoint_refuses_mixed_batch():
    insts = mixed_instances()[:2]
    with heom_amd.Handle.from_instances(insts) as h:
        with pytest.raises(heom_amd.HeomError, match="share"):
            h.rhs_adjoint()
