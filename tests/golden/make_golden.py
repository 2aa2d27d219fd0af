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

``--full-size`` writes ``full/excitonic8_d8_full.npz`` instead (kept apart from the small
fixtures, which the oracle re-runs in the CPU suite): the whole 8-site, 16-mode,
depth-8 workload (735,471 ADOs) for all 500 steps of its config, about 50 minutes of
oracle time on 8 cores.  It holds

* ``tops``         [51, 8, 8]     the top ADO every 10 steps (first = start)
* ``tuples``, ``ados``            every ADO of tier <= 2 at the end, by tuple
* ``norm_tuples``  [k, 16] int8   every ADO whose uid is a multiple of 61, by tuple
* ``norms``        [k]            their squared Frobenius norms at the end
* ``tier_norm_sum``, ``tier_norm_max``, ``tier_count``  [9]  per-tier sums and maxima
  of the squared norms over all ADOs, and the tier sizes
* ``norm_checksum``  sum over all ADOs of norm * (1 + (sum_m (m+1) n_m mod 13) / 13),
  a tuple-weighted sum that does not depend on the ADO order
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


def tuple_weights(tuples):
    """1 + (sum_m (m+1) n_m mod 13) / 13: a weight fixed by the tuple alone."""
    m = np.arange(1, tuples.shape[1] + 1, dtype=np.int64)
    return 1.0 + ((tuples.astype(np.int64) * m[None, :]).sum(axis=1) % 13) / 13.0


def make_full(steps=500, observe_steps=10, stride=61):
    path = os.path.join(CONFIGS, "excitonic8_d8.cfg")
    cfg = O.parse_config(path)
    inst = O.instance_from_config(cfg)
    y = inst.state_with_top(cfg.rho_init)
    tops = []
    for r in range(steps // observe_steps):
        t = inst.propagate(y, cfg.step_size, observe_steps, observe_steps)
        if r == 0:
            tops.append(t[0])
        tops.append(t[-1])
        print(f"  {(r + 1) * observe_steps} of {steps} steps", flush=True)
    tops = np.stack(tops)
    T = inst.graph.tuples
    tier = T.sum(axis=1)
    keep = np.nonzero(tier <= 2)[0]
    s = y.reshape(inst.matrices, inst.states, inst.states, 2)
    ados = s[keep, ..., 0] + 1j * s[keep, ..., 1]
    norms = inst.norms(y)
    sub = np.arange(0, inst.matrices, stride)
    D = int(tier.max())
    tier_sum = np.array([norms[tier == d].sum() for d in range(D + 1)])
    tier_max = np.array([norms[tier == d].max() for d in range(D + 1)])
    tier_count = np.array([int((tier == d).sum()) for d in range(D + 1)])
    checksum = float((norms * tuple_weights(T)).sum())
    full = os.path.join(os.path.dirname(OUT), "full")
    os.makedirs(full, exist_ok=True)
    np.savez_compressed(os.path.join(full, "excitonic8_d8_full.npz"),
                        tops=tops, tuples=T[keep].astype(np.int32), ados=ados,
                        norm_tuples=T[sub].astype(np.int8), norms=norms[sub],
                        tier_norm_sum=tier_sum, tier_norm_max=tier_max, tier_count=tier_count,
                        norm_checksum=np.array(checksum), config=np.array(os.path.relpath(path, ROOT)),
                        depth=np.array(cfg.ado_depth), steps=np.array(steps),
                        observe_steps=np.array(observe_steps), step_size=np.array(cfg.step_size),
                        mode=np.array("identity"))
    print(f"excitonic8_d8_full: {inst.matrices} ADOs, {len(keep)} kept, {len(sub)} norms, {tops.shape[0]} tops")


if __name__ == "__main__":
    O.build()
    os.makedirs(OUT, exist_ok=True)
    if "--full-size" in sys.argv[1:]:
        make_full()
    else:
        for case in CASES:
            make(*case)
