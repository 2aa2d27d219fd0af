"""Static-disorder ensembles on one GPU: every realisation of a system propagated in
one batch handle (heom_create_batch), population dynamics and linear absorption.

The reference runs its apps once per realisation (dm-heom/src/app_population_dynamics
.cpp, app_linear_absorption.cpp).  Here each realisation's config file goes through
the same host path (config -> hierarchy graph -> instance), and all realisations are
stepped together in one batch handle: the population hierarchies (the 7 sites, padded
to 8 states by the batch with a state that has no bath and stays zero) and the
absorption hierarchies (the sites plus the ground state, one instance per dipole tensor
component), one kernel launch per RK4 stage for the whole ensemble.  Populations and
absorption traces are reduced on the device.

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

        # instances 0 .. members-1: populations; then ncomp absorption instances per member
        self.handle = Handle._create(descs(self.pop_hosts, 1) + descs(self.la_hosts, self.ncomp),
                                     keepalive=self.pop_hosts + self.la_hosts)
        self.handle.use_graphs(use_graphs)
        self.pop_ados = sum(h.matrices for h in self.pop_hosts)
        self.la_ados = sum(h.matrices for h in self.la_hosts) * self.ncomp

        # rotating frame: each realisation's mean site energy leaves the excited block of
        # its absorption instances; the population instances keep their Hamiltonian
        Hs, self.e0 = [], []
        for hst in self.pop_hosts:
            H = hst.arrays()["hamiltonian"].reshape(hst.states, hst.states, 2)
            Hs.append(H[..., 0] + 1j * H[..., 1])
        for hst in self.la_hosts:
            H = hst.arrays()["hamiltonian"].reshape(hst.states, hst.states, 2)
            Hc = H[..., 0] + 1j * H[..., 1]
            e0 = float(np.mean(np.real(np.diag(Hc)[1:])))
            Hc = Hc.copy()
            Hc[np.arange(1, hst.states), np.arange(1, hst.states)] -= e0
            self.e0.append(e0)
            Hs.extend([Hc] * self.ncomp)
        self.handle.set_hamiltonian(Hs)
        self.dminus, self.dplus = [], []
        for hst in self.la_hosts:
            for t in range(self.ncomp):
                dm, dp = hst.dipoles(t)
                self.dminus.append(dm)
                self.dplus.append(dp)
        eye = [np.eye(h.states) for h in self.pop_hosts]
        zero = [np.zeros((h.states, h.states)) for h in self.pop_hosts]
        self.start_ops = eye + self.dplus    # D+ on the absorption instances, identity elsewhere
        self.trace_ops = zero + self.dminus  # Tr(D- sigma_0) on the absorption instances
        # the stage-kernel configuration timed on the batch itself (heom_autotune)
        self.tuning = {}
        if autotune:
            self.reset()
            self.tuning["ensemble"] = next(r for r in self.handle.autotune(steps=3) if r["chosen"])

    def reset(self) -> None:
        """Initial states: |site_1><site_1| (populations) and D+ |g><g| (absorption)."""
        self.handle.clear_state()
        self.handle.set_top([h.top() for h in self.pop_hosts] +
                            [h.top() for h in self.la_hosts for _ in range(self.ncomp)])
        self.handle.mmult_left(self.start_ops)

    def run(self, steps: int | None = None) -> dict:
        """Propagate the ensemble; populations (rows, members, sites) and the averaged
        absorption correlation C(t) (rows, members), with the device time."""
        steps = self.steps if steps is None else steps
        self.reset()
        pops, tr = self.handle.propagate_observe(self.step_size, steps, self.observe_steps, self.trace_ops)
        t = self.handle.timing().last_call_seconds
        rows = pops.shape[0]
        m, sites = self.members, self.pop_hosts[0].states
        times = np.arange(rows) * self.observe_steps * self.step_size
        phase = np.exp(-1j * np.outer(times, np.array(self.e0)) / HBAR)
        corr = tr[:, m:].reshape(rows, m, self.ncomp).mean(axis=2) * phase
        ados = self.pop_ados + self.la_ados
        return {
            "times": times, "populations": pops[:, :m, :sites], "absorption": corr, "steps": steps,
            "device_seconds": t, "ado_steps": ados * steps, "ados": ados,
        }

    def close(self) -> None:
        self.handle.close()
        for hst in self.pop_hosts + self.la_hosts:
            hst.close()
        self._tmp.cleanup()
