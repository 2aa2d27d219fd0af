"""Explicit Runge-Kutta steppers on the per-evaluation interface (heom_evaluate,
heom_lincomb, heom_swap_buffers of include/heom_hip.h).

The reference selects its stepper by name (`solver.stepper_type`,
dm-heom/src/heom/config.cpp:32-34 and :49) from noma/num's `stepper_type_names` and builds
it with `num::make_stepper` (dm-heom/include/heom/population_dynamics_solver.hpp:112-114);
noma/num is not vendored with the reference, so its list of names and its schemes are not
available here.  Its example configs use "rk_rk4" (classical RK4), which heom_step_rk4
runs fused.  This module composes explicit Runge-Kutta methods from their Butcher
tableaus (textbook coefficients: Butcher, Numerical Methods for Ordinary Differential
Equations, 2016, sec. 23-32) on the device buffers of an FP64 handle:

* a tableau whose only non-zero a_ij are a_{i,i-1} (Euler, midpoint, Heun, classical RK4)
  runs in the fused accumulating form of the reference's ODE wrapper: each evaluation
  writes the next stage input y + h a_{i+1,i} k_i and adds h b_i k_i to the accumulator
  (four buffers);
* any other explicit tableau (Kutta's third-order method, the 3/8 rule) keeps every k_i
  and forms the stage inputs and the result with heom_lincomb (s + 2 buffers).

The names other than "rk_rk4" are this module's.  `numpy_step` runs the same tableau on
the host with any linear right-hand side, for the tests.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import HEOM_NO_BUFFER


@dataclass(frozen=True)
class Tableau:
    a: tuple  # a[i] = (a_i0, ..., a_i,i-1) for stages i = 1 .. s-1
    b: tuple
    c: tuple
    order: int

    @property
    def stages(self) -> int:
        return len(self.b)

    @property
    def subdiagonal(self) -> bool:
        return all(v == 0.0 for row in self.a for v in row[:-1])


TABLEAUS = {
    "rk_euler": Tableau(a=(), b=(1.0,), c=(0.0,), order=1),
    "rk_midpoint": Tableau(a=((0.5,),), b=(0.0, 1.0), c=(0.0, 0.5), order=2),
    "rk_heun": Tableau(a=((1.0,),), b=(0.5, 0.5), c=(0.0, 1.0), order=2),
    "rk_kutta3": Tableau(a=((0.5,), (-1.0, 2.0)), b=(1 / 6, 2 / 3, 1 / 6), c=(0.0, 0.5, 1.0), order=3),
    "rk_rk4": Tableau(a=((0.5,), (0.0, 0.5), (0.0, 0.0, 1.0)), b=(1 / 6, 1 / 3, 1 / 3, 1 / 6),
                      c=(0.0, 0.5, 0.5, 1.0), order=4),
    "rk_rk4_38": Tableau(a=((1 / 3,), (-1 / 3, 1.0), (1.0, -1.0, 1.0)), b=(1 / 8, 3 / 8, 3 / 8, 1 / 8),
                         c=(0.0, 1 / 3, 2 / 3, 1.0), order=4),
}


def tableau(name_or_tableau) -> Tableau:
    if isinstance(name_or_tableau, Tableau):
        return name_or_tableau
    try:
        return TABLEAUS[name_or_tableau]
    except KeyError:
        raise ValueError(f"unknown stepper {name_or_tableau!r}; one of {', '.join(TABLEAUS)}") from None


class ExplicitRK:
    """An explicit Runge-Kutta method on a handle's state (buffer 0)."""

    def __init__(self, handle, method="rk_rk4"):
        self.h = handle
        self.t = tableau(method)
        if self.t.stages > 7:
            raise ValueError("at most 7 stages (a stage input combines y and the earlier k_i)")
        need = 4 if self.t.subdiagonal else self.t.stages + 2
        have = handle.buffer_count()
        if have < need:
            handle.add_buffers(need - have)

    def step(self, time: float, dt: float) -> None:
        """One step from `time` (passed to the pre-evaluate actions as time + c_i dt)."""
        t, h = self.t, self.h
        s = t.stages
        if t.subdiagonal:
            # buffer 1 accumulates y + dt sum b_i k_i; stage inputs alternate in 2 and 3
            src = 0
            for i in range(s):
                if i + 1 < s:
                    nxt = 2 + (i % 2)
                    h.evaluate(time + t.c[i] * dt, dt, src, out=nxt, out_base=0, out_coeff=dt * t.a[i][i],
                               acc=1, acc_coeff=dt * t.b[i], acc_init=(i == 0))
                    src = nxt
                else:
                    h.evaluate(time + t.c[i] * dt, dt, src, acc=1, acc_coeff=dt * t.b[i], acc_init=(i == 0))
            h.swap_buffers(0, 1)
            return
        # general explicit tableau: k_i in buffers 2 .. s+1, the stage input in buffer 1
        k = [2 + i for i in range(s)]
        for i in range(s):
            src = 0
            if i > 0:
                terms = [(1.0, 0)] + [(dt * a, k[j]) for j, a in enumerate(t.a[i - 1]) if a != 0.0]
                h.lincomb(1, terms)
                src = 1
            h.evaluate(time + t.c[i] * dt, dt, src, out=k[i])
        h.lincomb(1, [(1.0, 0)] + [(dt * b, k[i]) for i, b in enumerate(t.b) if b != 0.0])
        h.swap_buffers(0, 1)

    def propagate(self, time: float, dt: float, nsteps: int, observe_steps: int) -> np.ndarray:
        """nsteps steps; the tops [rows, count, N, N] every observe_steps steps (the first row
        the initial state), as Handle.propagate records them."""
        rows = [self.h.get_top()]
        for n in range(1, nsteps + 1):
            self.step(time + (n - 1) * dt, dt)
            if n % observe_steps == 0:
                rows.append(self.h.get_top())
        return np.stack(rows)


def numpy_step(rhs, y: np.ndarray, time: float, dt: float, method="rk_rk4", pre=None) -> np.ndarray:
    """The same tableau on the host: y_{n+1} from y_n with rhs(y) (and pre(time) before
    each evaluation)."""
    t = tableau(method)
    ks = []
    for i in range(t.stages):
        yi = y.copy()
        if i > 0:
            for j, a in enumerate(t.a[i - 1]):
                if a != 0.0:
                    yi = yi + dt * a * ks[j]
        if pre is not None:
            pre(time + t.c[i] * dt)
        ks.append(rhs(yi))
    out = y.copy()
    for i, b in enumerate(t.b):
        if b != 0.0:
            out = out + dt * b * ks[i]
    return out
