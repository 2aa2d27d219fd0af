"""Hierarchy filtering with compaction (heom_compact): the device layout keeps only the
unmasked ADOs.  Masked ADOs are never read by their neighbours, so the unmasked ones --
and every observable -- evolve exactly as in a masked handle without compaction and as
in the oracle with the same mask; the masked ones read back as zero."""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, config_path, instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


def random_mask(inst, seed, keep=0.6):
    rng = np.random.default_rng(seed)
    m = (rng.random(inst.matrices) < keep).astype(np.uint8)
    m[0] = O.MASK_PROTECTED
    return m


@pytest.mark.parametrize("name,depth", [("excitonic7_d4", None), ("excitonic8_d8", 3)])
def test_compacted_equals_masked_and_oracle(oracle_mod, name, depth):
    cfg, inst = instance(name, depth=depth)
    mask = random_mask(inst, 5)
    y0 = random_state(inst, 9)
    res = {}
    for compact in (False, True):
        with heom_amd.Handle.from_instance(inst) as h:
            h.set_mask(mask)
            if compact:
                kept = h.compact()
                assert kept == int(np.count_nonzero(mask))
                assert np.array_equal(h.get_mask(), mask)
            h.set_state(y0)
            tops = h.propagate(cfg.step_size, 60, 20)
            res[compact] = (tops, h.get_state(), h.rk4_step_bytes())
    keep = np.nonzero(mask)[0]
    a, b = as_complex(res[False][1], inst), as_complex(res[True][1], inst)
    assert max_abs_rel(res[True][0], res[False][0]) < 1e-13
    assert max_abs_rel(b[keep], a[keep]) < 1e-13
    drop = np.nonzero(mask == 0)[0]
    assert np.all(b[drop] == 0)
    assert res[True][2] * inst.matrices == res[False][2] * len(keep)
    inst.mask[:] = mask
    y = y0.copy()
    ref = inst.propagate(y, cfg.step_size, 60, 20)
    inst.mask[:] = O.MASK_TRUE
    err = max_abs_rel(res[True][0], ref)
    print(f"{name}: kept {len(keep)} of {inst.matrices} ADOs; tops vs the oracle {err:.2e}", flush=True)
    assert err < 1e-10
    assert max_abs_rel(b[keep], as_complex(y, inst)[keep]) < 1e-10


def test_single_excitation_filter(oracle_mod, tmp_path):
    """The reference's single_excitation strategy (hierarchy_mask.cpp:45-110), compacted."""
    src = open(config_path("excitonic7_d4")).read()
    src = src.replace("strategy=none", "strategy=single_excitation").replace("first_layer=-1", "first_layer=2")
    path = tmp_path / "se.cfg"
    path.write_text(src)
    host = heom_amd.HostInstance(str(path))
    mask = host.mask()
    cfg = O.parse_config(str(path))
    inst = O.instance_from_config(cfg)
    with host.create_handle() as h:
        h.set_mask(mask)
        kept = h.compact()
        tops = h.propagate(cfg.step_size, 200, 10)
    inst.mask[:] = mask
    ref = inst.propagate(inst.state_with_top(cfg.rho_init), cfg.step_size, 200, 10)
    err = max_abs_rel(tops, ref)
    print(f"single excitation, first layer 2: {kept} of {inst.matrices} ADOs, tops vs the oracle {err:.2e}",
          flush=True)
    assert kept < inst.matrices and err < 1e-10


def test_norm_threshold_then_compact_twice(oracle_mod):
    cfg, inst = instance("excitonic7_d4")
    y0 = inst.state_with_top(cfg.rho_init)
    with heom_amd.Handle.from_instance(inst) as h, heom_amd.Handle.from_instance(inst) as g:
        for x in (h, g):
            x.set_state(y0)
            x.step_rk4(cfg.step_size, 50)
        # thresholds at the 30th and 50th percentile of the current norms
        n1 = h.norms()
        t1 = float(np.quantile(n1, 0.3))
        h.norms(t1)
        g.norms(t1)
        assert np.array_equal(h.get_mask(), g.get_mask())
        k1 = h.compact()
        h.step_rk4(cfg.step_size, 50)
        g.step_rk4(cfg.step_size, 50)
        n2 = h.norms()
        t2 = float(np.quantile(n2[n2 > 0], 0.5))
        h.norms(t2)
        g.norms(t2)
        assert np.array_equal(h.get_mask(), g.get_mask())
        k2 = h.compact()
        assert k2 < k1 < inst.matrices
        th, tg = h.propagate(cfg.step_size, 40, 10), g.propagate(cfg.step_size, 40, 10)
        print(f"kept {k1}, then {k2} of {inst.matrices} ADOs", flush=True)
    assert max_abs_rel(th, tg) < 1e-13
    assert n1.size == inst.matrices


def test_compacted_batch(oracle_mod):
    _, a = instance("excitonic7_d4", depth=3)
    _, b = instance("excitonic7_d4", depth=4)
    ma, mb = random_mask(a, 1), random_mask(b, 2)
    top = np.zeros((7, 7), np.complex128)
    top[1, 1] = 1.0
    with heom_amd.Handle.from_instances([a, b]) as hb:
        hb.set_mask(np.concatenate([ma, mb]))
        hb.compact()
        hb.clear_state()
        hb.set_top(top)
        rec = hb.propagate(1e-15, 40, 10)
    for k, (inst, m) in enumerate(((a, ma), (b, mb))):
        with heom_amd.Handle.from_instance(inst) as h:
            h.set_mask(m)
            h.clear_state()
            h.set_top(top)
            rec1 = h.propagate(1e-15, 40, 10)
        assert max_abs_rel(rec[:, k], rec1) < 1e-13


def test_compact_rejects(oracle_mod):
    _, inst = instance("excitonic7_d4", depth=2)
    m = np.ones(inst.matrices, np.uint8)
    m[0] = 0
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_mask(m)
        with pytest.raises(heom_amd.HeomError, match="top"):
            h.compact()
    with heom_amd.Handle.from_instance(inst, precision=heom_amd.HEOM_PRECISION_FP32) as h:
        with pytest.raises(heom_amd.HeomError, match="FP64"):
            h.compact()


def test_adjoint_on_compacted_handle(oracle_mod):
    """After compaction the masked ADOs are gone, so the adjoint runs; its gradients
    equal the oracle's adjoint of the masked operator."""
    _, inst = instance("excitonic7_d4", depth=3)
    mask = random_mask(inst, 3)
    rng = np.random.default_rng(4)
    N = inst.states
    y0 = inst.state_with_top(np.diag(rng.random(N) / N))
    op = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    w = rng.normal(size=11) + 1j * rng.normal(size=11)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_mask(mask)
        h.compact()
        h.set_state(y0)
        X = h.adjoint_forward(1e-15, 30, 3, op, checkpoint_steps=7)[:, 0]
        g = h.adjoint_backward(w)
    inst.mask[:] = mask
    Xr, gr = inst.adjoint_gradient(y0, 1e-15, 30, 3, op, w, checkpoint_steps=7)
    inst.mask[:] = O.MASK_TRUE
    assert max_abs_rel(X, Xr) < 1e-12
    assert max_abs_rel(g["hamiltonian"][0], gr["hamiltonian"]) < 1e-10
    for k in ("pf_same", "pf_circ", "cbk"):
        assert max_abs_rel(g[k], gr[k]) < 1e-10, k
    keep = np.nonzero(mask)[0]
    assert max_abs_rel(g["pf_same_sum"][keep], gr["pf_same_sum"][keep]) < 1e-10
