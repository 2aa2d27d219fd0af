"""heom_observe: the populations and traces of the current state, reduced on the device,
against the tops read back (per instance of a batch, FP64 and FP32 storage) and against
the first row of heom_propagate_observe."""
import numpy as np
import pytest

from helpers import instance, random_state

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


@pytest.mark.parametrize("precision", [heom_amd.HEOM_PRECISION_FP64, heom_amd.HEOM_PRECISION_FP32])
def test_observe_matches_tops(oracle_mod, precision):
    _, a = instance("excitonic8_d6", depth=2)
    _, b = instance("excitonic8_d6", depth=3)
    rng = np.random.default_rng(1)
    ops = rng.normal(size=(2, 8, 8)) + 1j * rng.normal(size=(2, 8, 8))
    with heom_amd.Handle.from_instances([a, b], precision=precision) as h:
        h.set_state(np.concatenate([random_state(a, 1), random_state(b, 2)]))
        h.step_rk4(1e-15, 4)
        tops = h.get_top()
        pops, tr = h.observe(ops)
        pops_only, none = h.observe()
        rows_p, rows_t = h.propagate_observe(1e-15, 0, 1, ops)
    assert none is None
    want_p = np.real(np.diagonal(tops, axis1=1, axis2=2))
    want_t = np.einsum("kij,kji->k", ops, tops)
    assert np.allclose(pops, want_p, rtol=0, atol=1e-14 * np.abs(want_p).max())
    assert np.array_equal(pops_only, pops)
    assert np.allclose(tr, want_t, rtol=1e-13, atol=1e-14 * np.abs(want_t).max())
    assert np.array_equal(rows_p[0], pops) and np.array_equal(rows_t[0], tr)
