"""The static-disorder ensemble (heom_amd.ensemble): every realisation's populations and
rotating-frame absorption trace from the batched GPU run against the oracle run of the
same realisation's config files (1e-10 relative), and the FP32 ensemble against the
FP64 one (population error at most 1e-6 over the full 2048 steps of 2 fs)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

heom_amd = pytest.importorskip("heom_amd")
from heom_amd.ensemble import Ensemble  # noqa: E402


def test_ensemble_fp64_matches_oracle(oracle_mod):
    ens = Ensemble(members=5, steps=200, observe_steps=8)
    try:
        res = ens.run()
        for m in range(ens.members):
            cfg = O.parse_config(os.path.join(ens._tmp.name, f"pop{m}.cfg"))
            inst = O.instance_from_config(cfg)
            y = inst.state_with_top(cfg.rho_init)
            tops = inst.propagate(y, cfg.step_size, 200, 8)
            pops = np.real(np.diagonal(tops, axis1=1, axis2=2))
            err = np.max(np.abs(res["populations"][:, m] - pops)) / np.max(np.abs(pops))
            lcfg = O.parse_config(os.path.join(ens._tmp.name, f"la{m}.cfg"))
            _, ref = O.linear_absorption(lcfg, nsteps=200, observe_steps=8, rotating_frame=ens.e0[m])
            lerr = np.max(np.abs(res["absorption"][:, m] - ref)) / np.max(np.abs(ref))
            print(f"member {m} depth {ens.depths[m]}: populations {err:.2e}, absorption {lerr:.2e}", flush=True)
            assert err < 1e-10 and lerr < 1e-10
    finally:
        ens.close()


def test_ensemble_fp32_population_error():
    results = {}
    for prec in (heom_amd.HEOM_PRECISION_FP64, heom_amd.HEOM_PRECISION_FP32):
        ens = Ensemble(members=12, precision=prec)
        try:
            results[prec] = ens.run()
        finally:
            ens.close()
    a, b = results[heom_amd.HEOM_PRECISION_FP64], results[heom_amd.HEOM_PRECISION_FP32]
    perr = np.max(np.abs(a["populations"] - b["populations"]))
    aerr = np.max(np.abs(a["absorption"] - b["absorption"])) / np.max(np.abs(a["absorption"]))
    print(f"FP32 ensemble over {a['steps']} steps: population error {perr:.2e}, absorption {aerr:.2e}", flush=True)
    assert perr < 1e-6
    assert aerr < 1e-4
