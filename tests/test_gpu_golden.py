"""The product against the oracle's committed golden vectors (tests/golden/vectors), with
no oracle run: each config goes through the host library (config -> graph -> instance ->
heom_create), is propagated on the GPU, and its tops and the tier <= 2 ADOs, read back by
tuple (heom_get_ados), must match within 1e-10 relative to the largest value."""
import glob
import os

import numpy as np
import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")

FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "vectors", "*.npz")))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_product_matches_golden(path):
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    mode = str(g["mode"])
    task = heom_amd.TASK_POPULATION_DYNAMICS if mode == "identity" else heom_amd.TASK_LINEAR_ABSORPTION
    host = heom_amd.HostInstance(os.path.join(ROOT, str(g["config"])), task=task, depth=int(g["depth"]))
    with host.create_handle() as h:
        if mode != "identity":
            _, dp = host.dipoles(0)
            h.mmult_left(dp)  # sigma_0 = D+ |g><g|
        tops = h.propagate(float(g["step_size"]), int(g["steps"]), int(g["observe_steps"]))
        ados = h.get_ados(g["tuples"])
    e_top = np.max(np.abs(tops - g["tops"])) / np.max(np.abs(g["tops"]))
    e_ado = np.max(np.abs(ados - g["ados"])) / np.max(np.abs(g["ados"]))
    print(f"{os.path.basename(path)}: tops {e_top:.1e}, ADOs by tuple {e_ado:.1e}", flush=True)
    assert e_top < 1e-10 and e_ado < 1e-10
