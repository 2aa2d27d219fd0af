"""heom_propagate_observe_noalloc (observation records in the library's fixed device
scratch, no device allocation) against heom_propagate_observe on the same handles and
states: populations and traces bit for bit, for two calls in a row with the second on a
smaller input, in FP64 and FP32 storage, on a batch, with step graphs enabled; and a call
whose records do not fit the scratch is refused without touching the state."""
import numpy as np
import pytest

from helpers import instance

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")


def observable(N, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    return a + a.conj().T


def both(make, h, nsteps, observe_steps, ops):
    """(existing entry point, no-alloc entry point) from the same start state"""
    with make() as a:
        pa, ta = a.propagate_observe(h, nsteps, observe_steps, ops)
        sa = a.get_state()
    with make() as b:
        pb, tb = b.propagate_observe(h, nsteps, observe_steps, ops, noalloc=True)
        sb = b.get_state()
    return (pa, ta, sa), (pb, tb, sb)


@pytest.mark.parametrize("prec", [heom_amd.HEOM_PRECISION_FP64, heom_amd.HEOM_PRECISION_FP32])
def test_noalloc_matches_allocating_entry_point(prec):
    cfg, big = instance("excitonic7_d4")  # 330 ADOs
    _, small = instance("excitonic7_d4", depth=2)  # 36 ADOs
    runs = [(big, 100, 10), (small, 30, 5)]  # the second call has the smaller input
    for inst, nsteps, every in runs:
        y0 = inst.state_with_top(cfg.rho_init)

        def make():
            h = heom_amd.Handle.from_instance(inst, precision=prec)
            h.set_state(y0)
            return h

        ops = observable(inst.states, nsteps)
        (pa, ta, sa), (pb, tb, sb) = both(make, cfg.step_size, nsteps, every, ops)
        assert pa.shape == (nsteps // every + 1, 1, inst.states)
        assert np.array_equal(pa, pb) and np.array_equal(ta, tb) and np.array_equal(sa, sb)
        # populations only (no operator)
        with make() as h:
            p_only, none = h.propagate_observe(cfg.step_size, nsteps, every, noalloc=True)
        assert none is None and np.array_equal(p_only, pa)


def test_noalloc_batch_with_graphs():
    cfg, _ = instance("excitonic7_d4")
    insts = [instance("excitonic7_d4", depth=d)[1] for d in (3, 4, 2)]
    ops = [observable(7, k) for k in range(3)]
    rows = {}
    for noalloc in (False, True):
        with heom_amd.Handle.from_instances(insts) as h:
            h.set_top(np.stack([cfg.rho_init] * 3))
            h.use_graphs(True)
            rows[noalloc] = h.propagate_observe(cfg.step_size, 60, 6, ops, noalloc=noalloc)
            # a second, shorter call on the same handle continues from where it stopped
            rows[(noalloc, 2)] = h.propagate_observe(cfg.step_size, 8, 4, ops, noalloc=noalloc)
    for key in (False, (False, 2)):
        alt = True if key is False else (True, 2)
        assert np.array_equal(rows[key][0], rows[alt][0]) and np.array_equal(rows[key][1], rows[alt][1])


def test_noalloc_refuses_records_beyond_the_scratch():
    cfg, inst = instance("excitonic7_d4", depth=2)
    y0 = inst.state_with_top(cfg.rho_init)
    with heom_amd.Handle.from_instance(inst) as h:
        h.set_state(y0)
        # 8 MiB of scratch holds about 131k rows of 7 populations; ask for 200k rows
        with pytest.raises(heom_amd.HeomError, match="scratch"):
            h.propagate_observe(cfg.step_size, 200000, 1, noalloc=True)
        assert np.array_equal(h.get_state(), y0)
        p, _ = h.propagate_observe(cfg.step_size, 10, 1, noalloc=True)
        assert p.shape == (11, 1, 7)
