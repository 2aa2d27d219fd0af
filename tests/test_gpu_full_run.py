"""Run-length parity at full size: the 8-site, 16-mode, depth-8 hierarchy (735,471
complex128 ADOs) for all 500 steps of its config, on the GPU through the host library,
against the oracle's committed fixture (tests/golden/full/excitonic8_d8_full.npz, written
by tests/golden/make_golden.py --full-size; no oracle run here).

Both stage kernels that handle the hierarchy (the persistent pipelined kernel and the
one-workgroup-per-tile kernel) must give
  * the populations (diagonal of the top ADO, every 10 steps) and the whole top within
    1e-10 relative to its largest element;
  * every tier <= 2 ADO at the end, read back by tuple, within 1e-10;
  * the squared Frobenius norms of a tuple-keyed subset of 12,057 ADOs (every uid that is a
    multiple of 61) within 1e-10 relative to max(norm, 1e-12 * max norm), since the norms
    span many orders of magnitude;
  * the per-tier sums and maxima of the norms of all ADOs, and a tuple-weighted checksum
    of all norms that does not depend on the ADO order, within 1e-10 relative.
"""
import os

import numpy as np
import pytest

from helpers import ROOT, rel_err

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")

FIXTURE = os.path.join(ROOT, "tests", "golden", "full", "excitonic8_d8_full.npz")


def tuple_weights(tuples):
    """1 + (sum_m (m+1) n_m mod 13) / 13: the checksum weight, fixed by the tuple alone."""
    m = np.arange(1, tuples.shape[1] + 1, dtype=np.int64)
    return 1.0 + ((tuples.astype(np.int64) * m[None, :]).sum(axis=1) % 13) / 13.0


@pytest.fixture(scope="module")
def golden():
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def host(golden):
    h = heom_amd.HostInstance(os.path.join(ROOT, str(golden["config"])), depth=int(golden["depth"]))
    yield h
    h.close()


def run(host, golden, config):
    with host.create_handle() as h:
        if config is not None:
            h.set_stage_config(config)
        kind = h.stage_kernel()
        tops = h.propagate(float(golden["step_size"]), int(golden["steps"]), int(golden["observe_steps"]))
        ados = h.get_ados(golden["tuples"])
        norms = h.norms()
    return kind, tops, ados, norms


@pytest.mark.parametrize("kernel", ["pipelined", "tile"])
def test_full_size_500_steps_match_oracle(host, golden, kernel):
    assert host.matrices == 735471
    config = None if kernel == "pipelined" else {"kernel": "tile", "xcd_chunk": 32, "workgroups": 0,
                                                  "ados_per_workgroup": 64}
    kind, tops, ados, norms = run(host, golden, config)
    assert kind == kernel

    g_tops = golden["tops"]
    e_top = np.max(np.abs(tops - g_tops)) / np.max(np.abs(g_tops))
    pops, g_pops = np.real(np.diagonal(tops, axis1=1, axis2=2)), np.real(np.diagonal(g_tops, axis1=1, axis2=2))
    e_pop = np.max(np.abs(pops - g_pops)) / np.max(np.abs(g_pops))
    e_ado = np.max(np.abs(ados - golden["ados"])) / np.max(np.abs(golden["ados"]))

    # norms by tuple: the host's uid -> tuple table gives each uid's tuple
    W = int(golden["norm_tuples"].shape[1])
    tuples = host.arrays()["tuples"].reshape(-1, W)
    index = {t.tobytes(): u for u, t in enumerate(tuples.astype(np.int8))}
    sub = np.array([index[t.tobytes()] for t in golden["norm_tuples"].astype(np.int8)])
    e_norm = rel_err(norms[sub], golden["norms"])
    tier = tuples.sum(axis=1)
    D = int(tier.max())
    tier_sum = np.array([norms[tier == d].sum() for d in range(D + 1)])
    tier_max = np.array([norms[tier == d].max() for d in range(D + 1)])
    tier_count = np.array([int((tier == d).sum()) for d in range(D + 1)])
    assert np.array_equal(tier_count, golden["tier_count"])
    e_tsum = rel_err(tier_sum, golden["tier_norm_sum"])
    e_tmax = rel_err(tier_max, golden["tier_norm_max"])
    checksum = float((norms * tuple_weights(tuples)).sum())
    e_chk = abs(checksum - float(golden["norm_checksum"])) / abs(float(golden["norm_checksum"]))
    print(f"{kernel}: 500 steps, tops {e_top:.1e}, populations {e_pop:.1e}, tier<=2 ADOs {e_ado:.1e}, "
          f"norms by tuple {e_norm:.1e}, tier sums {e_tsum:.1e}, tier maxima {e_tmax:.1e}, checksum {e_chk:.1e}",
          flush=True)
    assert e_top < 1e-10 and e_pop < 1e-10 and e_ado < 1e-10
    assert e_norm < 1e-10 and e_tsum < 1e-10 and e_tmax < 1e-10 and e_chk < 1e-10
    assert abs(np.trace(tops[-1]).real - 1.0) < 1e-12
