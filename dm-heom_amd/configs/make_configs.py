#!/usr/bin/env python3
"""Seeded generator of the workload config files (reference INI format).

Workloads (all FP64 inputs; energies in cm^-1, times in fs, T in K):

* ``excitonic7_d4``  7 sites, diagonal U[12000, 12600], couplings U[-100, 100],
  one Drude-Lorentz bath per site (lambda 35, 1/nu 50 fs, T 300 K, one mode per bath),
  depth 4 (330 ADOs), RK4 dt = 1 fs, 1000 steps.                            seed 0
* ``excitonic8_d8``  8 sites, couplings U[-120, 120], per-site bath with the Drude
  term plus one Pade/Matsubara pole (K = 2, 16 modes), depth 8 (735,471 ADOs),
  dt = 1 fs, 500 steps.                                                      seed 1
* ``excitonic24_d4`` 24 sites, couplings U[-150, 150], 24 baths, K = 1, depth 4
  (20,475 ADOs), dt = 1 fs, 1000 steps.                                      seed 3
* ``excitonic8_d6``  8 sites, couplings U[-100, 100], 8 baths, K = 1, depth 6
  (3,003 ADOs), dt = 1 fs, 1000 steps (the adjoint-gradient system).         seed 4
* ``ensemble7``      512 static-disorder realisations of a 7-site H: couplings and
  base diagonal from seed 2, site energies + N(0, 50^2) per realisation from seed
  [2, i], depth drawn from {3, 4, 5, 6} per realisation, dt = 2 fs, 2048 steps;
  population dynamics plus linear absorption (random unit dipoles, xx/yy/zz).
  Generated on demand by ``ensemble_member(i)``; not written to disk.

RNG is numpy's default_rng (PCG64).  Draw order: diagonal (n values), then the
strict upper triangle row by row.  Values are rounded to 1e-6 cm^-1 when written.
The initial density matrix is |site_1><site_1| with 0-based site numbering (the
second site), as in the reference's output header ``site_0 ... site_{N-1}``.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def random_hamiltonian(rng, n, lo=12000.0, hi=12600.0, jmax=100.0):
    diag = rng.uniform(lo, hi, size=n)
    H = np.diag(diag)
    for i in range(n):
        for j in range(i + 1, n):
            v = rng.uniform(-jmax, jmax)
            H[i, j] = H[j, i] = v
    return np.round(H, 6)


def fmt_matrix(M, fmt="{:.6f}"):
    return "{" + ", ".join("{" + ", ".join(fmt.format(v) for v in row) + "}" for row in M) + "}"


def fmt_vector(v, fmt="{:g}"):
    return "{" + ", ".join(fmt.format(x) for x in v) + "}"


def population_cfg(name, H, depth, matsubaras, steps, step_fs, observe_steps=10,
                   lam=35.0, invnu=50.0, temperature=300.0, rho_site=1,
                   observations=None):
    n = H.shape[0]
    rho = np.zeros((n, n))
    rho[rho_site, rho_site] = 1.0
    obs = observations or [("matrix_diagonal", f"{name}_diagonal.dat")]
    lines = [
        "[program]",
        "task=population_dynamics",
        "observations={" + ", ".join(f"({a}, {b})" for a, b in obs) + "}",
        f"observe_steps={observe_steps}",
        "",
        "[filtering]",
        "strategy=none",
        "first_layer=-1",
        "",
        "[solver]",
        "stepper_type=rk_rk4",
        f"step_size={step_fs:g}e-15",
        f"steps={steps}",
        "track_flows=false",
        "flow_filename=",
        "",
        "[system]",
        f"sites={n}",
        f"hamiltonian={fmt_matrix(H)}",
        f"ado_depth={depth}",
        "",
        "[baths]",
        f"number={n}",
        "max_per_site=1",
        "coupling=" + "{" + ", ".join("{%d}" % i for i in range(n)) + "}",
        "lambda=" + fmt_vector([lam] * n),
        "invnu=" + fmt_vector([invnu] * n),
        "Omega=" + fmt_vector([0] * n),
        f"matsubaras={matsubaras}",
        f"temperature={temperature:g}",
        "",
        "[population_dynamics]",
        f"rho_init={fmt_matrix(rho, '{:.1f}')}",
        "",
    ]
    return "\n".join(lines)


def absorption_cfg(name, H, depth, matsubaras, steps, step_fs, dipoles, observe_steps=1,
                   lam=35.0, invnu=50.0, temperature=300.0):
    n = H.shape[0]
    lines = [
        "[program]",
        "task=linear_absorption",
        f"observations={{(matrix_trace_static_fluorescence, {name}_linear_absorption.dat)}}",
        f"observe_steps={observe_steps}",
        "",
        "[filtering]",
        "strategy=none",
        "first_layer=-1",
        "",
        "[solver]",
        "stepper_type=rk_rk4",
        f"step_size={step_fs:g}e-15",
        f"steps={steps}",
        "track_flows=false",
        "flow_filename=",
        "",
        "[system]",
        f"sites={n}",
        f"hamiltonian={fmt_matrix(H)}",
        f"ado_depth={depth}",
        "",
        "[baths]",
        f"number={n}",
        "max_per_site=1",
        "coupling=" + "{" + ", ".join("{%d}" % i for i in range(n)) + "}",
        "lambda=" + fmt_vector([lam] * n),
        "invnu=" + fmt_vector([invnu] * n),
        "Omega=" + fmt_vector([0] * n),
        f"matsubaras={matsubaras}",
        f"temperature={temperature:g}",
        "",
        "[dipole]",
        f"directions={fmt_matrix(dipoles, '{:.10f}')}",
        "centers=" + fmt_matrix(np.zeros((n, 3)), "{:.1f}"),
        "strengths=" + fmt_vector([1] * n),
        "tensor_prefactors={1, 1, 1}",
        "tensor_components={{0, 0}, {1, 1}, {2, 2}}",
        "",
    ]
    return "\n".join(lines)


def excitonic7_d4():
    H = random_hamiltonian(np.random.default_rng(0), 7, jmax=100.0)
    return population_cfg("excitonic7_d4", H, depth=4, matsubaras=1, steps=1000, step_fs=1)


def excitonic8_d8():
    H = random_hamiltonian(np.random.default_rng(1), 8, jmax=120.0)
    return population_cfg("excitonic8_d8", H, depth=8, matsubaras=2, steps=500, step_fs=1)


def excitonic24_d4():
    H = random_hamiltonian(np.random.default_rng(3), 24, jmax=150.0)
    return population_cfg("excitonic24_d4", H, depth=4, matsubaras=1, steps=1000, step_fs=1)


def excitonic8_d6():
    H = random_hamiltonian(np.random.default_rng(4), 8, jmax=100.0)
    return population_cfg("excitonic8_d6", H, depth=6, matsubaras=1, steps=1000, step_fs=1)


def ensemble_base():
    """Couplings, base diagonal and dipole directions shared by all realisations."""
    rng = np.random.default_rng(2)
    H0 = random_hamiltonian(rng, 7, jmax=100.0)
    dirs = rng.normal(size=(7, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return H0, dirs


def ensemble_member(i, H0=None, dirs=None):
    """(H, depth, dipoles) of realisation i of the 512-member ensemble."""
    if H0 is None:
        H0, dirs = ensemble_base()
    rng = np.random.default_rng([2, i])
    H = H0.copy()
    H[np.diag_indices(7)] += rng.normal(0.0, 50.0, size=7)
    depth = int(rng.integers(3, 7))  # {3, 4, 5, 6}: 120 .. 1716 ADOs
    return np.round(H, 6), depth, dirs


WRITERS = {
    "excitonic7_d4": excitonic7_d4,
    "excitonic8_d8": excitonic8_d8,
    "excitonic24_d4": excitonic24_d4,
    "excitonic8_d6": excitonic8_d6,
}


def path_of(name):
    return os.path.join(HERE, f"{name}.cfg")


def main(argv):
    for name, fn in WRITERS.items():
        with open(path_of(name), "w") as f:
            f.write(fn())
        print("wrote", path_of(name))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
