"""The stage-kernel autotuner (heom_autotune): it times every valid configuration on
the handle's own state, leaves that state untouched, keeps the fastest, and the kept
configuration still matches the oracle."""
import numpy as np
import pytest

from helpers import instance, max_abs_rel, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


@pytest.mark.parametrize("name,depth,prec", [
    ("excitonic8_d8", 3, "fp64"), ("excitonic7_d4", None, "fp64"), ("excitonic8_d8", 3, "fp32"),
    ("excitonic24_d4", 2, "fp64"), ("excitonic24_d4", 2, "fp32"),
])
def test_autotune_keeps_state_and_parity(oracle_mod, name, depth, prec):
    cfg, inst = instance(name, depth=depth)
    precision = heom_amd.HEOM_PRECISION_FP32 if prec == "fp32" else heom_amd.HEOM_PRECISION_FP64
    y = random_state(inst, 11)
    with heom_amd.Handle.from_instance(inst, precision=precision) as h:
        h.set_state(y)
        before = h.get_state()
        res = h.autotune(steps=2)
        assert np.array_equal(h.get_state(), before)
        chosen = [r for r in res if r["chosen"]]
        assert len(chosen) == 1
        assert chosen[0]["ms_per_step"] == min(r["ms_per_step"] for r in res)
        assert h.stage_kernel() == chosen[0]["kernel"]
        kinds = {r["kernel"] for r in res}
        if inst.states > 8:
            assert kinds == {"wide"}
        else:
            assert "tile" in kinds
        print(f"{name} {prec}: {len(res)} candidates, kept {chosen[0]}", flush=True)
        h.set_state(inst.state_with_top(cfg.rho_init))
        tops = h.propagate(cfg.step_size, 30, 10)
    ref = inst.propagate(inst.state_with_top(cfg.rho_init), cfg.step_size, 30, 10)
    assert max_abs_rel(tops, ref) < (1e-10 if prec == "fp64" else 1e-5)


def test_autotune_batch_with_graphs(oracle_mod):
    _, a = instance("excitonic7_d4", depth=3)
    _, b = instance("excitonic7_d4", depth=4)
    top = np.zeros((7, 7), np.complex128)
    top[1, 1] = 1.0
    recs = []
    for tune in (False, True):
        with heom_amd.Handle.from_instances([a, b]) as h:
            h.use_graphs(True)
            h.clear_state()
            h.set_top(top)
            h.propagate(1e-15, 20, 10)  # builds the step graph
            if tune:
                h.autotune(steps=2)
            recs.append(h.propagate(1e-15, 40, 10))
    assert max_abs_rel(recs[1], recs[0]) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pipelined_chunks_agree_bitwise(prec):
    """The XCD chunk only changes which workgroup takes which tile and when: every chunk
    the autotuner offers (powers of two and multiples of 24) and an odd one give the same
    bits after five RK4 steps, on a hierarchy of 318 tiles (not a multiple of 8 chunks)."""
    _, inst = instance("excitonic8_d8", depth=5)
    precision = heom_amd.HEOM_PRECISION_FP32 if prec == "fp32" else heom_amd.HEOM_PRECISION_FP64
    y = random_state(inst, 5)
    with heom_amd.Handle.from_instance(inst, precision=precision) as h:
        pipe = [r for r in h.autotune(steps=1) if r["kernel"] == "pipelined"]
        assert pipe and {24, 48, 72, 96} <= {r["xcd_chunk"] for r in pipe}
        out = {}
        for chunk in (1, 7, 24, 32, 48, 64, 72, 96):
            h.set_stage_config({**pipe[0], "xcd_chunk": chunk})
            h.set_state(y)
            h.step_rk4(1e-15, 5)
            out[chunk] = h.get_state()
    for chunk, v in out.items():
        assert np.array_equal(v, out[1]), chunk
