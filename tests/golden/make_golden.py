#!/usr/bin/env python3
"""Writes the oracle's golden vectors (tests/golden/vectors/*.npz), keyed by ADO tuple.

Each file holds one short propagation of a seeded workload config by the CPU oracle
(oracle/heom_oracle.c through oracle/oracle.py):

* ``tuples``  [k, W] int32    the ADOs kept, by tuple (every ADO of tier <= 2)
* ``ados``    [k, N, N]       those ADOs at the end of the run
* ``tops``    [rows, N, N]    the top ADO every ``observe_steps`` steps (first = start)
* ``config``, ``depth``, ``steps``, ``observe_steps``, ``step_size``, ``mode``

Cases: the 7-site depth-4 workload (330 ADOs, 200 steps), the 8-site depth-6 adjoint
system (3003 ADOs, 100 steps), the 8-site 16-mode system at depth 3 (969 ADOs, 50 steps)
and the FMO linear-absorption example with its ground state (depth 2, 100 steps of the
D+ |g><g| start).  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "vectors")
CONFIGS = os.path.join(ROOT, "dm-heom_amd", "configs")
FMO = os.path.join(ROOT, "tests", "golden", "fmo")

CASES = [
    # name, config file, depth override, steps, observe_steps, sites-to-states mode
    ("excitonic7_d4", os.path.join(CONFIGS, "excitonic7_d4.cfg"), None, 200, 20, "identity"),
    ("excitonic8_d6", os.path.join(CONFIGS, "excitonic8_d6.cfg"), None, 100, 10, "identity"),
    ("excitonic8_d8_depth3", os.path.join(CONFIGS, "excitonic8_d8.cfg"), 3, 50, 10, "identity"),
    ("fmo_linear_absorption_depth2", os.path.join(FMO, "fmo_linear_absorption_77K.cfg"), 2, 100, 10,
     "with_ground_state"),
]


def start(cfg, inst, mode):
    if mode == "identity":
        return inst.state_with_top(cfg.rho_init)
    dm, dp = O.dipole_matrices(cfg, 0)
    g = np.zeros((inst.states, inst.states), dtype=np.complex128)
    g[0, 0] = 1.0
    return inst.state_with_top(dp @ g)


def make(name, path, depth, steps, observe_steps, mode):
    cfg = O.parse_config(path)
    if depth is not None:
        cfg.ado_depth = depth
    inst = O.instance_from_config(cfg, mode)
    y = start(cfg, inst, mode)
    tops = inst.propagate(y, cfg.step_size, steps, observe_steps)
    keep = np.nonzero(inst.graph.tuples.sum(axis=1) <= 2)[0]
    s = y.reshape(inst.matrices, inst.states, inst.states, 2)
    ados = s[keep, ..., 0] + 1j * s[keep, ..., 1]
    np.savez_compressed(os.path.join(OUT, name + ".npz"), tuples=inst.graph.tuples[keep].astype(np.int32),
                        ados=ados, tops=tops, config=np.array(os.path.relpath(path, ROOT)),
                        depth=np.array(cfg.ado_depth), steps=np.array(steps), observe_steps=np.array(observe_steps),
                        step_size=np.array(cfg.step_size), mode=np.array(mode))
    print(f"{name}: {inst.matrices} ADOs, {len(keep)} kept, {tops.shape[0]} tops")


if __name__ == "__main__":
    O.build()
    os.makedirs(OUT, exist_ok=True)
    for case in CASES:
        make(*case)
