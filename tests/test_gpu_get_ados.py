"""Tuple-keyed readback (heom_get_ados): ADOs addressed by their tuples agree with the
uid-ordered state of heom_get_state for FP64 and FP32 storage (tier-0/1 ADOs from their
FP64 copies), per instance of a batch, after compaction (removed ADOs read as zero), and
tuples that are not ADOs of the instance are refused."""
import numpy as np
import pytest

import oracle as O
from helpers import as_complex, instance, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


@pytest.mark.parametrize("precision", [heom_amd.HEOM_PRECISION_FP64, heom_amd.HEOM_PRECISION_FP32])
def test_get_ados_matches_state(oracle_mod, precision):
    _, inst = instance("excitonic8_d6", depth=3)
    y = random_state(inst, 1)
    rng = np.random.default_rng(2)
    pick = np.concatenate([[0], rng.choice(inst.matrices, 40, replace=False)])
    with heom_amd.Handle.from_instance(inst, precision=precision) as h:
        h.set_state(y)
        h.step_rk4(1e-15, 3)
        state = as_complex(h.get_state(), inst)
        got = h.get_ados(inst.graph.tuples[pick])
    assert np.array_equal(got, state[pick])


def test_get_ados_per_instance_and_after_compaction(oracle_mod):
    _, a = instance("excitonic8_d6", depth=2)
    _, b = instance("excitonic8_d6", depth=3)
    ys = [random_state(a, 3), random_state(b, 4)]
    with heom_amd.Handle.from_instances([a, b]) as h:
        h.set_state(np.concatenate(ys))
        for k, inst in enumerate((a, b)):
            got = h.get_ados(inst.graph.tuples, instance=k)
            assert np.array_equal(got, as_complex(ys[k], inst))
    with heom_amd.Handle.from_instance(b) as h:
        h.set_state(ys[1])
        mask = np.ones(b.matrices, dtype=np.uint8)
        mask[np.random.default_rng(5).choice(np.arange(1, b.matrices), 30, replace=False)] = 0
        h.set_mask(mask)
        h.compact()
        got = h.get_ados(b.graph.tuples)
    want = as_complex(ys[1], b)
    want[mask == 0] = 0
    assert np.array_equal(got, want)


def test_get_ados_refuses_foreign_tuples(oracle_mod):
    _, inst = instance("excitonic8_d6", depth=2)
    with heom_amd.Handle.from_instance(inst) as h:
        deep = np.zeros((1, inst.width), dtype=np.int32)
        deep[0, 0] = 3  # depth 3 > 2
        for bad in (deep, -np.ones((1, inst.width), dtype=np.int32)):
            with pytest.raises(heom_amd.HeomError, match="not an ADO"):
                h.get_ados(bad)
        with pytest.raises(heom_amd.HeomError, match="instance out of range"):
            h.get_ados(inst.graph.tuples[:1], instance=1)
        assert h.get_ados(np.zeros((0, inst.width), dtype=np.int32)).shape == (0, inst.states, inst.states)
