"""Differentiable HEOM observables for fitting: torch.autograd over the adjoint of the
RK4 trajectory.

``System(config)`` holds a reference config file's system (population dynamics, or
linear absorption with its ground state).  ``System.observe(H, lam, gamma, op)``
returns X_r = Tr(O sigma_0(t_r)) as a complex128 torch tensor that is differentiable
with respect to the site Hamiltonian H (sites x sites) and the bath reorganisation
energies lambda_b and rates gamma_b = hbar / invnu_b, all in cm^-1.  The forward pass
is heom_adjoint_forward (propagation on the GPU with checkpoints every
``checkpoint_steps`` steps); backward runs heom_adjoint_backward -- the transposed
operator integrated backwards by the adjoint kernels, contracted with the recomputed
forward states -- and the host chain rule heom_host_parameter_gradient.  Any loss
built from X with torch (time integrals, Fourier transforms into spectra, ...) can then
be minimised over H, lambda and gamma.

The reference has no differentiation; its config grammar and host pipeline are the
ones this uses (dm-heom/src/heom/config.cpp, dm-heom/src/heom/instance.cpp).
"""
from __future__ import annotations

import numpy as np
import torch

from . import TASK_POPULATION_DYNAMICS, HostInstance


class _Observe(torch.autograd.Function):
    @staticmethod
    def forward(ctx, hamiltonian, lam, gamma, system, op, nsteps, step_size, observe_steps):
        params = (hamiltonian.detach().cpu().numpy(), lam.detach().cpu().numpy(), gamma.detach().cpu().numpy())
        handle = system._handle(*params)
        try:
            X = handle.adjoint_forward(step_size, nsteps, observe_steps, op, system.checkpoint_steps)[:, 0]
        except BaseException:
            handle.close()
            raise
        ctx.system, ctx.handle, ctx.params = system, handle, params
        ctx.complex_h = torch.is_complex(hamiltonian)
        return torch.from_numpy(X)

    @staticmethod
    def backward(ctx, grad_x):
        handle = ctx.handle
        try:
            # J = sum_r Re(w_r X_r) has dJ/dX_r = conj(w_r) in torch's convention
            w = np.conj(grad_x.detach().cpu().numpy().astype(np.complex128))
            g = handle.adjoint_backward(w, state=False)
        finally:
            handle.close()
            ctx.handle = None
        host = ctx.system.host
        host.set_parameters(*ctx.params)  # the chain rule is taken at the forward's parameters
        pg = host.parameter_gradient(g)
        dH = pg["hamiltonian"] if ctx.complex_h else pg["hamiltonian"].real
        return (torch.from_numpy(np.ascontiguousarray(dH)), torch.from_numpy(pg["lambda"]),
                torch.from_numpy(pg["gamma"]), None, None, None, None, None)


class System:
    """A config file's system with its Hamiltonian and bath parameters as inputs."""

    def __init__(self, config_path: str, task: int = TASK_POPULATION_DYNAMICS, depth: int = 0,
                 checkpoint_steps: int = 50, use_graphs: bool = True, prepare=None):
        self.host = HostInstance(config_path, task, depth)
        self.checkpoint_steps = int(checkpoint_steps)
        self.use_graphs = use_graphs
        self.prepare = prepare  # called with the new handle before the forward sweep
        info = self.host.info
        self.step_size = float(info.step_size)
        self.steps = int(info.steps)
        self.states = self.host.states
        self.sites = self.host.sites
        p = self.host.get_parameters()
        self._initial = p

    def parameters(self, requires_grad: bool = True) -> dict:
        """The config's H (real part), lambda and gamma as float64 leaf tensors (cm^-1)."""
        p = self._initial
        return {k: torch.tensor(np.real(v), dtype=torch.float64, requires_grad=requires_grad)
                for k, v in (("hamiltonian", p["hamiltonian"]), ("lam", p["lambda"]), ("gamma", p["gamma"]))}

    def _handle(self, hamiltonian, lam, gamma):
        self.host.set_parameters(hamiltonian, lam, gamma)
        handle = self.host.create_handle()
        handle.use_graphs(self.use_graphs)
        if self.prepare is not None:
            self.prepare(handle)
        return handle

    def observe(self, hamiltonian, lam, gamma, op, nsteps: int | None = None, step_size: float | None = None,
                observe_steps: int = 1) -> torch.Tensor:
        """X_r = Tr(op sigma_0(t_r)), r = 0 .. nsteps / observe_steps (complex128, rows)."""
        nsteps = self.steps if nsteps is None else int(nsteps)
        step_size = self.step_size if step_size is None else float(step_size)
        op = np.asarray(op, dtype=np.complex128)
        return _Observe.apply(hamiltonian, lam, gamma, self, op, nsteps, step_size, int(observe_steps))

    def population_integral(self, site: int, hamiltonian, lam, gamma, nsteps: int | None = None,
                            step_size: float | None = None) -> torch.Tensor:
        """J = integral of the population of `site` (0-based) over [0, nsteps * step_size]
        by the trapezoidal rule on every RK4 step, in fs."""
        nsteps = self.steps if nsteps is None else int(nsteps)
        step_size = self.step_size if step_size is None else float(step_size)
        op = np.zeros((self.states, self.states), dtype=np.complex128)
        s = site + self.states - self.sites
        op[s, s] = 1.0
        X = self.observe(hamiltonian, lam, gamma, op, nsteps, step_size, 1).real
        return step_size * 1e15 * (X.sum() - 0.5 * (X[0] + X[-1]))

    def close(self) -> None:
        self.host.close()
