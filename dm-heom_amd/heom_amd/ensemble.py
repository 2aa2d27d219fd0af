"""Static-disorder ensembles on one GPU: every realisation of a system propagated in
one batch handle (heom_create_batch), population dynamics and linear absorption.

The reference runs its apps once per realisation (dm-heom/src/app_population_dynamics
.cpp, app_linear_absorption.cpp).  Here each realisation's config file goes through
the same host path (config -> hierarchy graph -> instance), and all realisations are
stepped together: one batch of population hierarchies (the sites) and one batch of
absorption hierarchies (the sites plus the ground state, one instance per dipole
tensor component).  Populations and absorption traces are reduced on the device.

Linear absorption at a 2 fs step needs the rotating frame: the optical coherences
rotate at |E| (about 12,000 cm^-1, omega dt = 4.5), beyond classical RK4's stability
bound of 2.83, so each realisation propagates with its mean site energy E0 removed from
the excited block and the trace is multiplied by exp(-i E0 t / hbar), as in
heom_linear_absorption --rotating-frame.
"""
from __future__ import annotations

import importlib.util
import os
import tempfile
import time

import numpy as np

from . import (CONFIG_DIR, HEOM_PRECISION_FP64, TASK_LINEAR_ABSORPTION, TASK_POPULATION_DYNAMICS, Handle,
               HostInstance)

HBAR = 1.05457148e-34  # J s, dm-heom/include/heom/constants.hpp:21


def _make_configs():
    spec = importlib.util.spec_from_file_location("heom_make_configs", os.path.join(CONFIG_DIR, "make_configs.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Ensemble:
    """Realisations 0..members-1 of the seeded 7-site static-disorder ensemble
    (configs/make_configs.py: ensemble_member), each with population dynamics from
    |site_1><site_1| and linear absorption over the xx, yy, zz tensor components."""

    def __init__(self, members: int = 512, precision: int = HEOM_PRECISION_FP64, step_fs: float = 2.0,
                 steps: int = 2048, observe_steps: int = 8, use_graphs: bool = True, autotune: bool = True):
        mc = _make_configs()
        H0, dirs = mc.ensemble_base()
        self.members = members
        self.steps, self.observe_steps = steps, observe_steps
        self.step_size = step_fs * 1e-15
        self._tmp = tempfile.TemporaryDirectory(prefix="heom_ensemble_")
        t0 = time.perf_counter()
        self.pop_hosts, self.la_hosts, self.depths = [], [], []
        for i in range(members):
            H, depth, dip = mc.ensemble_member(i, H0, dirs)
            pop = os.path.join(self._tmp.name, f"pop{i}.cfg")
            la = os.path.join(self._tmp.name, f"la{i}.cfg")
            with open(pop, "w") as f:
                f.write(mc.population_cfg(f"m{i}", H, depth=depth, matsubaras=1, steps=steps, step_fs=step_fs,
                                          observe_steps=observe_steps))
            with open(la, "w") as f:
                f.write(mc.absorption_cfg(f"m{i}", H, depth=depth, matsubaras=1, steps=steps, step_fs=step_fs,
                                          dipoles=dip, observe_steps=observe_steps))
            self.pop_hosts.append(HostInstance(pop, TASK_POPULATION_DYNAMICS))
            self.la_hosts.append(HostInstance(la, TASK_LINEAR_ABSORPTION))
            self.depths.append(depth)
        self.setup_seconds = time.perf_counter() - t0
        self.ncomp = self.la_hosts[0].info.tensor_components

        def descs(hosts, repeat):
            out = []
            for hst in hosts:
                d = hst.desc()
                d.precision = int(precision)
                out.extend([d] * repeat)
            return out

        self.pop = Handle._create(descs(self.pop_hosts, 1), keepalive=self.pop_hosts)
        self.la = Handle._create(descs(self.la_hosts, self.ncomp), keepalive=self.la_hosts)
        for h in (self.pop, self.la):
            h.use_graphs(use_graphs)
        self.pop_ados = sum(h.matrices for h in self.pop_hosts)
        self.la_ados = sum(h.matrices for h in self.la_hosts) * self.ncomp

        # rotating frame: each realisation's mean site energy leaves the excited block
        Hs, self.e0 = [], []
        for hst in self.la_hosts:
            H = hst.arrays()["hamiltonian"].reshape(hst.states, hst.states, 2)
            Hc = H[..., 0] + 1j * H[..., 1]
            e0 = float(np.mean(np.real(np.diag(Hc)[1:])))
            Hc = Hc.copy()
            Hc[np.arange(1, hst.states), np.arange(1, hst.states)] -= e0
            self.e0.append(e0)
            Hs.extend([Hc] * self.ncomp)
        self.la.set_hamiltonian(np.array(Hs))
        self.dminus, self.dplus = [], []
        for hst in self.la_hosts:
            for t in range(self.ncomp):
                dm, dp = hst.dipoles(t)
                self.dminus.append(dm)
                self.dplus.append(dp)
        self.dminus, self.dplus = np.array(self.dminus), np.array(self.dplus)
        # stage-kernel configurations timed on the batches themselves (heom_autotune)
        self.tuning = {}
        if autotune:
            self.reset()
            for name, h in (("populations", self.pop), ("absorption", self.la)):
                self.tuning[name] = next(r for r in h.autotune(steps=3) if r["chosen"])

    def reset(self) -> None:
        """Initial states: |site_1><site_1| (populations) and D+ |g><g| (absorption)."""
        self.pop.clear_state()
        self.pop.set_top(np.array([h.top() for h in self.pop_hosts]))
        self.la.clear_state()
        self.la.set_top(np.array([h.top() for h in self.la_hosts for _ in range(self.ncomp)]))
        self.la.mmult_left(self.dplus)

    def run(self, steps: int | None = None) -> dict:
        """Propagate both batches; populations (rows, members, sites) and the averaged
        absorption correlation C(t) (rows, members), with device times."""
        steps = self.steps if steps is None else steps
        self.reset()
        pops, _ = self.pop.propagate_observe(self.step_size, steps, self.observe_steps)
        t_pop = self.pop.timing().last_call_seconds
        _, tr = self.la.propagate_observe(self.step_size, steps, self.observe_steps, self.dminus)
        t_la = self.la.timing().last_call_seconds
        rows = pops.shape[0]
        times = np.arange(rows) * self.observe_steps * self.step_size
        phase = np.exp(-1j * np.outer(times, np.array(self.e0)) / HBAR)
        corr = tr.reshape(rows, self.members, self.ncomp).mean(axis=2) * phase
        ados = self.pop_ados + self.la_ados
        return {
            "times": times, "populations": pops, "absorption": corr, "steps": steps,
            "device_seconds": t_pop + t_la, "pop_seconds": t_pop, "la_seconds": t_la,
            "ado_steps": ados * steps, "ados": ados,
        }

    def close(self) -> None:
        for h in (self.pop, self.la):
            h.close()
        for hst in self.pop_hosts + self.la_hosts:
            hst.close()
        self._tmp.cleanup()
