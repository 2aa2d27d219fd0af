"""The thermal state search on the GPU (heom_create_tss / heom_tss_step) against the
oracle's oracle_tss_step on the same inputs: f_TSS (heom_rhs on a search handle), single
and repeated iterations for the fused transform (N <= 8) and the four-pass one (N = 24),
a filtered hierarchy, batches, one iteration at 735,471 ADOs, the full search of the FMO
thermal-state config to convergence, and the entry points a search handle refuses."""
import numpy as np
import pytest

import oracle as O
from helpers import FMO_DIR, instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")

TSS_CFG = FMO_DIR + "/fmo_thermal_state_77K.cfg"


def search_case(name=None, depth=None):
    if name is None:
        cfg = O.parse_config(TSS_CFG)
        inst = O.instance_from_config(cfg)
    else:
        cfg, inst = instance(name, depth=depth)
    s = O.thermal_state_setup(inst.hamiltonian_matrix(), cfg.baths_temperature)
    inst.set_hamiltonian(s["pseudo_hamiltonian"])
    return cfg, inst, s


CASES = [(None, None), ("excitonic8_d6", 3), ("excitonic7_d4", None), ("excitonic24_d4", 1)]


@pytest.mark.parametrize("name,depth", CASES)
def test_rhs_tss_matches_oracle(oracle_mod, name, depth):
    _, inst, s = search_case(name, depth)
    y = random_state(inst, 1)
    with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
        h.set_state(y)
        got = h.rhs()
    err = max_abs_rel(got, inst.rhs_tss(y))
    print(f"{name or 'fmo_thermal_state'}: f_TSS max rel deviation {err:.1e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("name,depth", CASES)
def test_tss_steps_match_oracle(oracle_mod, name, depth):
    _, inst, s = search_case(name, depth)
    y = random_state(inst, 2)
    ref = y.copy()
    with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
        h.set_state(y)
        h.tss_step(1)
        one = h.get_state()
        h.tss_step(9)
        ten = h.get_state()
    inst.tss_step(ref, s, 1)
    e1 = max_abs_rel(one, ref)
    inst.tss_step(ref, s, 9)
    e10 = max_abs_rel(ten, ref)
    print(f"{name or 'fmo_thermal_state'} ({inst.matrices} ADOs, N={inst.states}): 1 step {e1:.1e}, "
          f"10 steps {e10:.1e}", flush=True)
    assert e1 < 1e-10 and e10 < 1e-10


def test_masked_search_matches_oracle(oracle_mod):
    _, inst, s = search_case("excitonic8_d6", 3)
    rng = np.random.default_rng(3)
    mask = np.where(rng.random(inst.matrices) < 0.3, O.MASK_FALSE, O.MASK_TRUE).astype(np.uint8)
    mask[0] = O.MASK_PROTECTED
    y = random_state(inst, 4)
    with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
        h.set_mask(mask)
        h.set_state(y)
        h.tss_step(5)
        got = h.get_state()
    inst.mask[:] = mask
    inst.tss_step(y, s, 5)
    keep = np.repeat(mask != O.MASK_FALSE, inst.states * inst.states * 2)
    assert max_abs_rel(got[keep], y[keep]) < 1e-10


def test_fmo_search_to_convergence(oracle_mod):
    """The reference loop (app_thermal_state_search.cpp:84-128) on the device, against the
    oracle's: same iteration count, same converged top."""
    cfg = O.parse_config(TSS_CFG)
    tops_ref, y_ref, it_ref = O.thermal_state_search(cfg)
    host = heom_amd.HostInstance(TSS_CFG, task=heom_amd.TASK_THERMAL_STATE_SEARCH)
    with host.create_tss_handle() as h:
        last = h.get_top()
        assert np.max(np.abs(last - tops_ref[0])) < 1e-14
        it = 0
        min_it = cfg.steps // cfg.observe_steps
        while it < cfg.tss_max_steps // cfg.observe_steps:
            h.tss_step(cfg.observe_steps)
            it += 1
            top = h.get_top()
            done = it > min_it and abs(np.abs(last).sum() - np.abs(top).sum()) < cfg.tss_delta
            last = top
            if done:
                break
        y = h.get_state()
    print(f"FMO thermal state: {it} iterations (oracle {it_ref}), top deviation "
          f"{np.max(np.abs(top - tops_ref[-1])):.1e}", flush=True)
    assert it == it_ref
    assert np.max(np.abs(top - tops_ref[-1])) < 1e-10
    assert max_abs_rel(y, y_ref) < 1e-8


def test_search_handle_refuses_propagation(oracle_mod):
    _, inst, s = search_case("excitonic8_d6", 2)
    with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
        for call in (lambda: h.step_rk4(1e-15, 1), lambda: h.propagate(1e-15, 2, 1),
                     lambda: h.set_hamiltonian(np.eye(inst.states)), lambda: h.compact()):
            with pytest.raises(heom_amd.HeomError, match="thermal-state-search"):
                call()
    with heom_amd.Handle.from_instance(inst) as h:
        with pytest.raises(heom_amd.HeomError, match="not a thermal-state-search handle"):
            h.tss_step(1)


@pytest.mark.parametrize("name,depth", [(None, None), ("excitonic24_d4", 1)])
def test_complex_transform_matches_oracle(oracle_mod, name, depth):
    """The transform for any complex C (the real-C kernel serves real symmetric H; the
    general one is what the reference's kernels apply to any C and eigenvalues)."""
    _, inst, s = search_case(name, depth)
    N = inst.states
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    cs = dict(s, c=q, eigenvalues=s["eigenvalues"] + 1j * rng.normal(size=N) * 1e-22)
    y = random_state(inst, 6)
    ref = y.copy()
    with heom_amd.Handle.thermal_state_search_from_instance(inst, cs) as h:
        h.set_state(y)
        h.tss_step(3)
        got = h.get_state()
    inst.tss_step(ref, cs, 3)
    err = max_abs_rel(got, ref)
    print(f"{name or 'fmo_thermal_state'}: complex C, 3 steps {err:.1e}", flush=True)
    assert err < 1e-10


def test_batched_search_equals_single_handles_and_oracle(oracle_mod):
    """Three static-disorder-like instances (different H, depths 2-4) in one search handle:
    every instance's state equals its own single-instance search bit for bit and the
    oracle's to 1e-10, for the fused (N = 8) and the four-pass (N = 24) transform."""
    for name, depths in (("excitonic8_d6", (2, 3, 4)), ("excitonic24_d4", (1, 2, 1))):
        cases = []
        for k, D in enumerate(depths):
            cfg, inst = instance(name, depth=D)
            H = inst.hamiltonian_matrix()
            H = H + np.diag(np.random.default_rng(k).normal(0, 50, inst.states)) * O.INVCM_TO_JOULE
            s = O.thermal_state_setup(H, cfg.baths_temperature)
            inst.set_hamiltonian(s["pseudo_hamiltonian"])
            cases.append((inst, s, random_state(inst, 10 + k)))
        with heom_amd.Handle.thermal_state_search_from_instances([c[0] for c in cases],
                                                                 [c[1] for c in cases]) as hb:
            hb.set_state(np.concatenate([c[2] for c in cases]))
            hb.tss_step(4)
            sb = hb.get_state()
        off = 0
        for inst, s, y in cases:
            with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
                h.set_state(y)
                h.tss_step(4)
                one = h.get_state()
            ref = y.copy()
            inst.tss_step(ref, s, 4)
            n = inst.state_size
            assert np.array_equal(sb[off:off + n], one), name
            assert max_abs_rel(one, ref) < 1e-10, name
            off += n


def test_masked_batched_search_equals_masked_single_handles(oracle_mod):
    """A mask over a batched search handle (the instances' masks concatenated) acts on
    each instance as on its own handle: bitwise the single masked searches, and the
    oracle's masked search to 1e-10 on the ADOs the mask keeps (as for one instance)."""
    cases = []
    for k, D in enumerate((3, 2)):
        cfg, inst = instance("excitonic8_d6", depth=D)
        s = O.thermal_state_setup(inst.hamiltonian_matrix(), cfg.baths_temperature)
        inst.set_hamiltonian(s["pseudo_hamiltonian"])
        rng = np.random.default_rng(20 + k)
        mask = np.where(rng.random(inst.matrices) < 0.3, O.MASK_FALSE, O.MASK_TRUE).astype(np.uint8)
        mask[0] = O.MASK_PROTECTED
        cases.append((inst, s, mask, random_state(inst, 30 + k)))
    with heom_amd.Handle.thermal_state_search_from_instances([c[0] for c in cases], [c[1] for c in cases]) as hb:
        hb.set_mask(np.concatenate([c[2] for c in cases]))
        hb.set_state(np.concatenate([c[3] for c in cases]))
        hb.tss_step(3)
        sb = hb.get_state()
    off = 0
    for inst, s, mask, y in cases:
        with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
            h.set_mask(mask)
            h.set_state(y)
            h.tss_step(3)
            one = h.get_state()
        ref = y.copy()
        inst.mask[:] = mask
        inst.tss_step(ref, s, 3)
        n = inst.state_size
        keep = np.repeat(mask != O.MASK_FALSE, inst.states * inst.states * 2)
        assert np.array_equal(sb[off:off + n], one)
        assert max_abs_rel(one[keep], ref[keep]) < 1e-10
        off += n


def test_full_size_iteration_matches_oracle(oracle_mod):
    """735,471 ADOs (8 states, 16 modes, depth 8): one iteration from a random hierarchy
    against the oracle, then 5 from the Boltzmann start, whose top keeps trace one and
    stays Hermitian."""
    cfg, inst, s = search_case("excitonic8_d8", None)
    y = random_state(inst, 40)
    ref = y.copy()
    inst.tss_step(ref, s, 1)
    with heom_amd.Handle.thermal_state_search_from_instance(inst, s) as h:
        h.set_state(y)
        h.tss_step(1)
        got = h.get_state()
        err = max_abs_rel(got, ref)
        print(f"excitonic8_d8 ({inst.matrices} ADOs): 1 iteration {err:.1e}", flush=True)
        assert err < 1e-10
        h.set_state(inst.state_with_top(s["top"]))
        h.tss_step(5)
        top = h.get_top()
    assert abs(np.trace(top) - 1.0) < 1e-12
    assert np.max(np.abs(top - top.conj().T)) < 1e-12
