"""Batches (heom_create_batch): several hierarchies with the same states and baths but
different Hamiltonians and depths, stepped together by one launch per RK4 stage.

Every ADO of a batch goes through exactly the arithmetic it gets in a handle of its
own, so batch and single-instance results must agree bit for bit; each instance is
also checked against the oracle (1e-10 relative).
"""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


def instance(N, depth, seed, K=1, lam=35.0):
    rng = np.random.default_rng(seed)
    H = np.diag(rng.uniform(12000, 12600, N)).astype(np.complex128)
    for i in range(N):
        for j in range(i + 1, N):
            H[i, j] = H[j, i] = rng.uniform(-100, 100)
    H *= O.INVCM_TO_JOULE
    g = O.build_graph(N * K, depth)
    return O.Instance(states=N, matsubaras=K, baths=N, coupling=np.arange(N, dtype=np.int32).reshape(N, 1),
                      hamiltonian=H, graph=g, lambda_=np.full(N, lam) * O.INVCM_TO_JOULE,
                      invnu=np.full(N, 50e-15), omega=np.zeros(N), temperature=300.0)


def batch_instances(N=7):
    return [instance(N, d, seed) for d, seed in ((3, 11), (2, 12), (4, 13), (3, 14))]


@pytest.mark.parametrize("N", [7, 4])
def test_batch_equals_single_handles_and_oracle(oracle_mod, N):
    insts = batch_instances(N)
    dt, steps = 1e-15, 60
    tops0 = []
    for k, inst in enumerate(insts):
        top = np.zeros((N, N), np.complex128)
        top[k % N, k % N] = 1.0
        tops0.append(top)
    with heom_amd.Handle.from_instances(insts) as hb:
        assert hb.count == len(insts)
        hb.clear_state()
        hb.set_top(np.array(tops0))
        rec = hb.propagate(dt, steps, 20)
        state_b = hb.get_state()
        top_b = hb.get_top()
    assert rec.shape == (steps // 20 + 1, len(insts), N, N)
    offset = 0
    for k, inst in enumerate(insts):
        with heom_amd.Handle.from_instance(inst) as h:
            h.clear_state()
            h.set_top(tops0[k])
            rec1 = h.propagate(dt, steps, 20)
            state1 = h.get_state()
        n = inst.state_size
        assert np.array_equal(rec[:, k], rec1), k
        assert np.array_equal(state_b[offset:offset + n], state1), k
        assert np.array_equal(top_b[k], rec1[-1]), k
        offset += n
        y = inst.state_with_top(tops0[k])
        ref = inst.propagate(y, dt, steps, 20)
        err = max_abs_rel(rec1, ref)
        print(f"instance {k}: M={inst.matrices} max rel deviation from the oracle {err:.2e}", flush=True)
        assert err < 1e-10
    assert offset == state_b.size


def test_batch_rhs_state_mask_and_mmult(oracle_mod):
    insts = batch_instances()
    N = insts[0].states
    ys = [random_state(inst, 40 + k) for k, inst in enumerate(insts)]
    rng = np.random.default_rng(5)
    mxs = rng.normal(size=(len(insts), N, N)) + 1j * rng.normal(size=(len(insts), N, N))
    with heom_amd.Handle.from_instances(insts) as hb:
        hb.set_state(np.concatenate(ys))
        assert np.array_equal(hb.get_state(), np.concatenate(ys))
        f_b = hb.rhs()
        norms_b = hb.norms()
        hb.mmult_left(mxs)
        mm_b = hb.get_state()
        masks = [np.ones(inst.matrices, np.uint8) for inst in insts]
        masks[1][::3] = 0
        hb.set_mask(np.concatenate(masks))
        assert np.array_equal(hb.get_mask(), np.concatenate(masks))
        hb.set_state(np.concatenate(ys))
        fm_b = hb.rhs()
    off = 0
    offm = 0
    for k, inst in enumerate(insts):
        n = inst.state_size
        ref = inst.rhs(ys[k])
        assert max_abs_rel(f_b[off:off + n], ref) < 1e-12
        assert np.allclose(norms_b[offm:offm + inst.matrices], inst.norms(ys[k]), rtol=1e-13, atol=0)
        y = ys[k].copy()
        inst.mmult_left(mxs[k], y)
        assert max_abs_rel(mm_b[off:off + n], y) < 1e-13
        inst.mask[:] = masks[k]
        refm = inst.rhs(ys[k])
        inst.mask[:] = O.MASK_TRUE
        assert max_abs_rel(fm_b[off:off + n], refm) < 1e-12
        off += n
        offm += inst.matrices


def test_batch_rejects_mixed_precision(oracle_mod):
    """Instances may differ in baths and states (tests/test_gpu_batch_mixed.py), not in
    storage precision."""
    a = instance(7, 3, 1)
    b = instance(7, 3, 2, lam=40.0)
    descs = [heom_amd._make_desc(precision=p, **heom_amd._instance_arrays(i))
             for i, p in ((a, heom_amd.HEOM_PRECISION_FP64), (b, heom_amd.HEOM_PRECISION_FP32))]
    with pytest.raises(heom_amd.HeomError, match="precision"):
        heom_amd.Handle._create([d for d, _ in descs], [k for _, k in descs])


def test_graph_replays_match_eager(oracle_mod):
    """One RK4 step captured as a HIP graph, replayed many times with eager work
    (state uploads, readbacks) in between, gives the eager results bit for bit."""
    insts = batch_instances()
    N = insts[0].states
    top = np.zeros((len(insts), N, N), np.complex128)
    top[:, 1, 1] = 1.0
    out = {}
    for graphs in (False, True):
        with heom_amd.Handle.from_instances(insts) as h:
            h.use_graphs(graphs)
            h.clear_state()
            h.set_top(top)
            recs = [h.propagate(1e-15, 30, 10)]
            s = h.get_state()
            h.set_state(s)
            recs.append(h.propagate(1e-15, 30, 10))
            h.step_rk4(1e-15, 7)
            recs.append(h.get_top())
            h.step_rk4(2e-15, 5)  # a new step size rebuilds the graph
            recs.append(h.get_top())
            out[graphs] = recs
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b)
