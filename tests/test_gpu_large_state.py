"""A state larger than 2 GiB: the 8-site, 16-mode hierarchy at depth 10 (5,311,735 ADOs,
5.4 GB per complex128 state buffer, 2.7 GB with complex64 planes).

The persistent kernels address every element plane through its own buffer resource, so
they run at this size (a single resource over the whole state would limit it to 2 GiB).
Three RK4 steps from |site_1><site_1| with the persistent kernel (the pipelined kernel in
FP64, the paired-plane kernel in FP32) and with the tile kernel, which indexes with
64-bit pointers; the tops and the ADO norms of the two must agree: 1e-12 relative in
FP64 (the same arithmetic in another order), 1e-5 in FP32.  Each kernel is checked
against the oracle at small sizes by the parity tests.
"""
import os

import numpy as np
import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")

DEPTH = 10
ADOS = 5_311_735  # C(16 + 10, 10)


@pytest.fixture(scope="module")
def host():
    h = heom_amd.HostInstance(os.path.join(heom_amd.CONFIG_DIR, "excitonic8_d8.cfg"), depth=DEPTH)
    assert h.matrices == ADOS
    yield h
    h.close()


def run(host, precision, kernel, monkeypatch):
    monkeypatch.setenv("HEOM_STAGE_KERNEL", kernel)
    with host.create_handle(precision=precision) as h:
        kind = h.stage_kernel()
        h.clear_state()
        top = np.zeros((host.states, host.states), dtype=np.complex128)
        top[1, 1] = 1.0
        h.set_top(top)
        h.step_rk4(host.info.step_size, 3)
        return kind, h.get_top(), h.norms()


@pytest.mark.parametrize("precision,tol", [(heom_amd.HEOM_PRECISION_FP64, 1e-12),
                                           (heom_amd.HEOM_PRECISION_FP32, 1e-5)])
def test_state_above_2gib(host, precision, tol, monkeypatch):
    kind, top_p, norms_p = run(host, precision, "pipelined", monkeypatch)
    assert kind == "pipelined"
    kind, top_t, norms_t = run(host, precision, "tile", monkeypatch)
    assert kind == "tile"
    e_top = np.max(np.abs(top_p - top_t)) / np.max(np.abs(top_t))
    scale = np.maximum(norms_t, 1e-12 * norms_t.max())
    e_norm = np.max(np.abs(norms_p - norms_t) / scale)
    print(f"precision {precision}: {ADOS} ADOs, top {e_top:.1e}, norms {e_norm:.1e}", flush=True)
    assert np.count_nonzero(norms_t) > ADOS // 100  # three steps reach deep into the hierarchy
    assert e_top < tol and e_norm < tol
